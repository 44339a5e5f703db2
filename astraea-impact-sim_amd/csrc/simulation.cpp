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
// Bodies come from the hydrostatic planet builder (astraea_build_planet) and are relaxed on the
// device (astraea_relax), both bodies at once, anchored to their profile temperatures.  Render
// frames (astraea_pack_render) and crust tracer records (astraea_extract_tracers) come from the
// device as in simulation.cpp:1805-1867.
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

// Host particle arrays in input order.
struct Particles {
    std::vector<float> x, y, z, vx, vy, vz, m, h, h0, rho, u, eps;
    std::vector<uint8_t> mat, body;

    size_t size() const { return m.size(); }
    template <class F>
    void each(F&& f) {
        for (auto* v : {&x, &y, &z, &vx, &vy, &vz, &m, &h, &h0, &rho, &u, &eps}) f(*v);
        f(mat);
        f(body);
    }
    void resize(size_t n) {
        each([n](auto& v) { v.resize(n); });
    }
    void append(const Particles& o) {
        Particles& t = *this;
        auto cat = [](auto& a, const auto& b) { a.insert(a.end(), b.begin(), b.end()); };
        cat(t.x, o.x); cat(t.y, o.y); cat(t.z, o.z); cat(t.vx, o.vx); cat(t.vy, o.vy); cat(t.vz, o.vz);
        cat(t.m, o.m); cat(t.h, o.h); cat(t.h0, o.h0); cat(t.rho, o.rho); cat(t.u, o.u); cat(t.eps, o.eps);
        cat(t.mat, o.mat); cat(t.body, o.body);
    }
    Particles subset(const std::vector<uint32_t>& idx) const {
        Particles o;
        o.resize(idx.size());
        for (size_t k = 0; k < idx.size(); ++k) {
            const uint32_t i = idx[k];
            o.x[k] = x[i]; o.y[k] = y[i]; o.z[k] = z[i]; o.vx[k] = vx[i]; o.vy[k] = vy[i]; o.vz[k] = vz[i];
            o.m[k] = m[i]; o.h[k] = h[i]; o.h0[k] = h0[i]; o.rho[k] = rho[i]; o.u[k] = u[i]; o.eps[k] = eps[i];
            o.mat[k] = mat[i]; o.body[k] = body[i];
        }
        return o;
    }
    // mass-weighted centre and mean velocity
    void moments(double c[3], double v[3]) const {
        double M = 0.0;
        for (int a = 0; a < 3; ++a) c[a] = v[a] = 0.0;
        for (size_t i = 0; i < size(); ++i) {
            const double mi = m[i];
            M += mi;
            c[0] += mi * x[i]; c[1] += mi * y[i]; c[2] += mi * z[i];
            v[0] += mi * vx[i]; v[1] += mi * vy[i]; v[2] += mi * vz[i];
        }
        if (M > 0.0)
            for (int a = 0; a < 3; ++a) { c[a] /= M; v[a] /= M; }
    }
    void shift(const double c[3], const double v[3]) {
        for (size_t i = 0; i < size(); ++i) {
            x[i] = (float)((double)x[i] - c[0]); y[i] = (float)((double)y[i] - c[1]); z[i] = (float)((double)z[i] - c[2]);
            vx[i] = (float)((double)vx[i] - v[0]); vy[i] = (float)((double)vy[i] - v[1]); vz[i] = (float)((double)vz[i] - v[2]);
        }
    }
    // centerAndZeroBulkVelocity (simulation.cpp:298-330)
    void centre() {
        double c[3], v[3];
        moments(c, v);
        shift(c, v);
    }
};

struct Body {
    Particles p;
    std::vector<float> temp;  // profile temperature of each particle (K), the relaxation anchor
    // crust tracers, per particle (empty when the body has none)
    std::vector<int32_t> tracer;
    std::vector<float> restR, restN;
    std::vector<uint8_t> ocean;
    uint32_t tracerCount = 0;
    double radiusPsu = 0.0, massKg = 0.0;
};

// LivingTerra crust tracers of a body about its centre (planet_builder.cpp:745-906).
void selectTracers(Body& b, uint32_t gridRes) {
    const size_t n = b.p.size();
    b.tracer.assign(n, -1);
    b.restR.assign(n, 0.0f);
    b.restN.assign(3 * n, 0.0f);
    b.ocean.assign(n, 0);
    check(astraea_select_tracers((uint32_t)n, b.p.x.data(), b.p.y.data(), b.p.z.data(), b.p.h0.data(), b.p.mat.data(),
                                 gridRes, b.tracer.data(), b.restR.data(), b.restN.data(), b.ocean.data(),
                                 &b.tracerCount),
          "astraea_select_tracers");
}

void uploadParticles(astraea_ctx* ctx, const Particles& s, const std::vector<uint32_t>& counts) {
    check(astraea_set_layout(ctx, (uint32_t)counts.size(), counts.data()), "astraea_set_layout");
    const std::pair<int, const void*> f[] = {
        {ASTRAEA_F_POS_X, s.x.data()}, {ASTRAEA_F_POS_Y, s.y.data()}, {ASTRAEA_F_POS_Z, s.z.data()},
        {ASTRAEA_F_VEL_X, s.vx.data()}, {ASTRAEA_F_VEL_Y, s.vy.data()}, {ASTRAEA_F_VEL_Z, s.vz.data()},
        {ASTRAEA_F_MASS, s.m.data()}, {ASTRAEA_F_H, s.h.data()}, {ASTRAEA_F_H0, s.h0.data()},
        {ASTRAEA_F_RHO, s.rho.data()}, {ASTRAEA_F_U, s.u.data()}, {ASTRAEA_F_EPS_FAIL, s.eps.data()},
        {ASTRAEA_F_MATERIAL, s.mat.data()}, {ASTRAEA_F_BODY, s.body.data()}};
    for (const auto& q : f) check(astraea_upload_field(ctx, q.first, q.second), "astraea_upload_field");
}

// The loaded particles in input (original id) order.
Particles downloadParticles(astraea_ctx* ctx, size_t n) {
    Particles raw;
    raw.resize(n);
    std::vector<uint32_t> oid(n);
    const std::pair<int, void*> f[] = {
        {ASTRAEA_F_POS_X, raw.x.data()}, {ASTRAEA_F_POS_Y, raw.y.data()}, {ASTRAEA_F_POS_Z, raw.z.data()},
        {ASTRAEA_F_VEL_X, raw.vx.data()}, {ASTRAEA_F_VEL_Y, raw.vy.data()}, {ASTRAEA_F_VEL_Z, raw.vz.data()},
        {ASTRAEA_F_MASS, raw.m.data()}, {ASTRAEA_F_H, raw.h.data()}, {ASTRAEA_F_H0, raw.h0.data()},
        {ASTRAEA_F_RHO, raw.rho.data()}, {ASTRAEA_F_U, raw.u.data()}, {ASTRAEA_F_EPS_FAIL, raw.eps.data()},
        {ASTRAEA_F_MATERIAL, raw.mat.data()}, {ASTRAEA_F_BODY, raw.body.data()}, {ASTRAEA_F_ORIGINAL_ID, oid.data()}};
    for (const auto& q : f) check(astraea_download_field(ctx, q.first, q.second), "astraea_download_field");
    std::vector<uint32_t> slot(n);
    for (size_t i = 0; i < n; ++i) {
        if (oid[i] >= n) throw std::runtime_error("downloadParticles: original id out of range");
        slot[oid[i]] = (uint32_t)i;
    }
    return raw.subset(slot);
}

// Rigid rotation about +Y at omega rad/s (applyRigidSpinAroundY, simulation.cpp:331-341).
void spinAroundY(Particles& p, double omegaRadS) {
    const double om = omegaRadS * units::T_UNIT_S;
    for (size_t i = 0; i < p.size(); ++i) {
        p.vx[i] = (float)(om * (double)p.z[i]);
        p.vy[i] = 0.0f;
        p.vz[i] = (float)(-om * (double)p.x[i]);
    }
}

// PlanetBuilder::buildPlanet (planet_builder.cpp:474-735) on the device library's hydrostatic
// builder: the archetype's layers at their mass fractions, the spec's surface and core
// temperatures, 1 bar at a gas giant's surface (planet_builder.cpp:407).  Bodies start at rest
// about their centre; the spin is applied after any relaxation.
Body buildBody(const BodyInitSpec& spec, uint32_t n, uint64_t seed) {
    if (!(spec.massKg > 0.0)) throw std::invalid_argument("BodyInitSpec::massKg must be positive");
    const ArchetypeComposition& comp = getArchetypeComposition(spec.archetype);
    astraea_planet_spec ps;
    astraea_default_planet_spec(&ps);
    ps.mass_kg = spec.massKg;
    ps.num_layers = 0;
    for (uint32_t k = 0; k < comp.layerCount; ++k) {
        const PlanetLayerDef& l = comp.layers[k];
        if (l.massFraction <= 0.0) continue;
        ps.layer_material[ps.num_layers] = (uint8_t)l.material;
        ps.layer_mass_frac[ps.num_layers] = l.massFraction;
        ++ps.num_layers;
    }
    ps.surface_temp_k = spec.surfaceTempK;
    ps.core_temp_k = spec.coreTempK;
    ps.surface_pressure_pa = spec.archetype == PlanetArchetype::GasGiant ? 1.0e5 : 0.0;
    ps.seed = seed;
    Body b;
    Particles& q = b.p;
    q.resize(n);
    b.temp.resize(n);
    astraea_body_out out{q.x.data(), q.y.data(), q.z.data(), q.m.data(), q.h.data(), q.rho.data(), q.u.data(),
                         q.eps.data(), q.mat.data()};
    astraea_planet_info info{};
    check(astraea_build_planet(&ps, n, &out, b.temp.data(), &info), "astraea_build_planet");
    q.h0 = q.h;
    std::fill(q.vx.begin(), q.vx.end(), 0.0f);
    std::fill(q.vy.begin(), q.vy.end(), 0.0f);
    std::fill(q.vz.begin(), q.vz.end(), 0.0f);
    std::fill(q.body.begin(), q.body.end(), (uint8_t)0);
    b.radiusPsu = info.radius_m / units::L_UNIT_M;
    b.massKg = info.mass_kg;
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
    // centres of the last telemetry (PSU), for render frames
    double sysCom[3] = {0.0, 0.0, 0.0}, livingCom[3] = {0.0, 0.0, 0.0};
    // crust tracers of the loaded state, by tracer index
    uint32_t tracerCount = 0, tracerGridRes = 0;
    std::vector<uint32_t> trParticle;
    std::vector<float> trRest;  // 4 per tracer
    std::vector<uint8_t> trOcean;
    uint32_t count = 0;
    bool livingBodyPresent = false;
    TelemetrySnapshot lastSnapshot{};
    std::vector<TelemetrySnapshot> history;

    explicit Impl(uint32_t maxP) : maxParticles(std::max(maxP, 1u)) {
        check(astraea_create(&ctx, 0, maxParticles, 2), "astraea_create");
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
        check(astraea_create(&ctx, 0, maxParticles, 2), "astraea_create");
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
        s.tracerCount = tracerCount;
        s.tracerGridRes = tracerCount > 0 ? tracerGridRes : 0;
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
        for (int a = 0; a < 3; ++a) {  // simulation.cpp:1033-1053
            sysCom[a] = total > 0.0 ? sys[a] : 0.0;
            livingCom[a] = m0 > 0.0 ? com0[a] : sysCom[a];
        }
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

    void setDefaultParams() {
        astraea_params p;
        astraea_default_params(&p);
        check(astraea_set_params(ctx, &p), "astraea_set_params");
    }

    // Damped relaxation of bodies at rest (loadScenario, simulation.cpp:1329-1351 and 1366-1385):
    // the bodies with the same step count relax together on the device, one segment each,
    // anchored to the builder's temperatures.  Copy-back as
    // copyRelaxedBodyToBuiltHost (simulation.cpp:783-846): positions, density and u of the
    // relaxed body, velocities zero, h = h0, the radius the largest |x|, then centred.
    void relaxBodies(const std::vector<std::pair<Body*, uint32_t>>& list) {
        std::vector<bool> done(list.size(), false);
        for (size_t a = 0; a < list.size(); ++a) {
            if (done[a] || list[a].second == 0) continue;
            std::vector<Body*> group;
            for (size_t k = a; k < list.size(); ++k)
                if (!done[k] && list[k].second == list[a].second) {
                    group.push_back(list[k].first);
                    done[k] = true;
                }
            Particles all;
            std::vector<uint32_t> counts;
            std::vector<double> radii;
            std::vector<float> anchor;
            for (Body* b : group) {
                anchor.insert(anchor.end(), b->temp.begin(), b->temp.end());
                std::fill(b->p.vx.begin(), b->p.vx.end(), 0.0f);
                std::fill(b->p.vy.begin(), b->p.vy.end(), 0.0f);
                std::fill(b->p.vz.begin(), b->p.vz.end(), 0.0f);
                b->p.centre();
                all.append(b->p);
                counts.push_back((uint32_t)b->p.size());
                radii.push_back(b->radiusPsu);
            }
            setDefaultParams();
            uploadParticles(ctx, all, counts);
            {  // the bodies' tracers stay flagged while they relax (the living-body cooling test)
                std::vector<uint32_t> id;
                std::vector<float> rest;
                std::vector<uint8_t> oc;
                uint32_t base = 0;
                for (Body* b : group) {
                    std::vector<std::pair<int32_t, uint32_t>> byT;
                    for (size_t i = 0; i < b->tracer.size(); ++i)
                        if (b->tracer[i] >= 0) byT.push_back({b->tracer[i], (uint32_t)i});
                    std::sort(byT.begin(), byT.end());
                    for (const auto& q : byT) {
                        id.push_back(base + q.second);
                        rest.push_back(b->restR[q.second]);
                        for (int a = 0; a < 3; ++a) rest.push_back(b->restN[3 * q.second + a]);
                        oc.push_back(b->ocean[q.second]);
                    }
                    base += (uint32_t)b->p.size();
                }
                check(astraea_set_tracers(ctx, (uint32_t)id.size(), id.data(), rest.data(), oc.data()),
                      "astraea_set_tracers");
            }
            std::vector<double> vr(group.size());
            check(astraea_relax(ctx, list[a].second, anchor.data(), radii.data(), vr.data()), "astraea_relax");
            const Particles r = downloadParticles(ctx, all.size());
            size_t off = 0;
            for (Body* b : group) {
                Particles& q = b->p;
                double rMax = 0.0;
                for (size_t i = 0; i < q.size(); ++i) {
                    q.x[i] = r.x[off + i]; q.y[i] = r.y[off + i]; q.z[i] = r.z[off + i];
                    q.rho[i] = r.rho[off + i]; q.u[i] = r.u[off + i];
                    q.vx[i] = q.vy[i] = q.vz[i] = 0.0f;
                    q.h[i] = q.h0[i];
                    const double ri = std::sqrt((double)q.x[i] * q.x[i] + (double)q.y[i] * q.y[i] + (double)q.z[i] * q.z[i]);
                    rMax = std::max(rMax, ri);
                    if (!b->tracer.empty() && b->tracer[i] >= 0) {  // tracers at rest where the body settled
                        b->restR[i] = (float)ri;
                        if (ri > 1.0e-9) {
                            b->restN[3 * i] = (float)(q.x[i] / ri);
                            b->restN[3 * i + 1] = (float)(q.y[i] / ri);
                            b->restN[3 * i + 2] = (float)(q.z[i] / ri);
                        }
                    }
                }
                if (rMax > 0.0) b->radiusPsu = rMax;
                q.centre();
                off += q.size();
            }
        }
    }

    // Forces, counters and history of a freshly uploaded state (simulation.cpp:1449-1453).
    // The tracer table on the device (cleared by every upload of a layout).
    void uploadTracers() {
        check(astraea_set_tracers(ctx, tracerCount, trParticle.data(), trRest.data(), trOcean.data()),
              "astraea_set_tracers");
    }

    void tracersFromBody(const Body& b, uint32_t firstParticle, uint32_t gridRes) {
        tracerCount = b.tracerCount;
        tracerGridRes = gridRes;
        trParticle.assign(tracerCount, 0);
        trRest.assign(4 * (size_t)tracerCount, 0.0f);
        trOcean.assign(tracerCount, 0);
        for (size_t i = 0; i < b.tracer.size(); ++i) {
            const int32_t t = b.tracer[i];
            if (t < 0 || (uint32_t)t >= tracerCount) continue;
            trParticle[t] = firstParticle + (uint32_t)i;
            trRest[4 * t] = b.restR[i];
            for (int a = 0; a < 3; ++a) trRest[4 * t + 1 + a] = b.restN[3 * i + a];
            trOcean[t] = b.ocean[i];
        }
    }

    void startFromUpload(uint32_t n) {
        uploadTracers();
        count = n;
        frameSeq = 0;
        simTimeSec = 0.0;
        lastDtSec = 0.0f;
        hasImpacted = false;
        history.clear();
        check(astraea_evaluate_initial_forces(ctx), "astraea_evaluate_initial_forces");
        check(astraea_synchronize(ctx), "astraea_synchronize");
        telemetry();
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
    if (config.target.archetype == PlanetArchetype::LivingTerra && config.tracerGridRes > 0)
        selectTracers(t, config.tracerGridRes);
    Body b;
    if (hasImpactor) b = buildBody(config.impactor, nI, config.randomSeed + 1ULL);
    std::vector<std::pair<Body*, uint32_t>> relax;
    if (config.target.relaxSteps > 0u) relax.push_back({&t, config.target.relaxSteps});
    if (hasImpactor && config.impactor.relaxSteps > 0u) relax.push_back({&b, config.impactor.relaxSteps});
    I.relaxBodies(relax);
    I.targetRadiusPsu = t.radiusPsu;
    I.livingBodyPresent = config.target.archetype == PlanetArchetype::LivingTerra;
    spinAroundY(t.p, config.target.spinOmegaRadS);
    Particles all = t.p;
    I.qRatioQrd = 0.0;
    if (hasImpactor) {
        I.impactorRadiusPsu = b.radiusPsu;
        spinAroundY(b.p, config.impactor.spinOmegaRadS);
        std::fill(b.p.body.begin(), b.p.body.end(), (uint8_t)1);
        all.append(b.p);
        double vEscPsu = 0.0;
        check(astraea_place_two_bodies(nT, nI, all.x.data(), all.y.data(), all.z.data(), all.vx.data(), all.vy.data(),
                                       all.vz.data(), all.m.data(), t.radiusPsu, b.radiusPsu,
                                       config.impactVelocityRatioVesc, config.impactParamB, config.initialSepRadii,
                                       &vEscPsu),
              "astraea_place_two_bodies");
        I.qRatioQrd = qRatio(t.massKg, b.massKg, t.radiusPsu * units::L_UNIT_M, b.radiusPsu * units::L_UNIT_M,
                             std::max(0.0, config.impactVelocityRatioVesc) * vEscPsu * units::V_UNIT_MS,
                             config.impactParamB);
    }
    I.tracersFromBody(t, 0, config.tracerGridRes);  // the target's particles come first
    I.setDefaultParams();
    uploadParticles(I.ctx, all, {(uint32_t)all.size()});
    I.startFromUpload((uint32_t)all.size());
}

void Simulation::relaxBody(uint8_t bodyId, uint32_t steps, double& outFinalVrmsOverVesc) {
    Impl& I = *impl_;
    if (I.count == 0u) throw std::runtime_error("Simulation::relaxBody called before loadScenario");
    // simulation.cpp:1464-1709: the body alone, centred, at rest when steps > 0; relaxed; put back
    // at its centre with its bulk velocity when other bodies exist; stress and damage cleared.
    const Particles all = downloadParticles(I.ctx, I.count);
    std::vector<uint32_t> sel, rest;
    for (uint32_t i = 0; i < I.count; ++i) (all.body[i] == bodyId ? sel : rest).push_back(i);
    if (sel.empty()) throw std::invalid_argument("Simulation::relaxBody: no particles for requested bodyId");
    Particles p = all.subset(sel);
    double c[3], v[3];
    p.moments(c, v);
    const double zero[3] = {0.0, 0.0, 0.0};
    p.shift(c, zero);
    if (steps > 0u) {
        std::fill(p.vx.begin(), p.vx.end(), 0.0f);
        std::fill(p.vy.begin(), p.vy.end(), 0.0f);
        std::fill(p.vz.begin(), p.vz.end(), 0.0f);
    }
    p.h = p.h0;
    double nominal = bodyId == 0u ? I.targetRadiusPsu : I.impactorRadiusPsu;
    I.setDefaultParams();
    uploadParticles(I.ctx, p, {(uint32_t)p.size()});
    std::vector<uint32_t> newIndex(I.count, 0);
    for (size_t k = 0; k < sel.size(); ++k) newIndex[sel[k]] = (uint32_t)k;
    for (size_t k = 0; k < rest.size(); ++k) newIndex[rest[k]] = (uint32_t)(sel.size() + k);
    {  // the body's tracers stay flagged while it relaxes
        std::vector<uint32_t> id;
        std::vector<float> tr;
        std::vector<uint8_t> oc;
        for (uint32_t t = 0; t < I.tracerCount; ++t) {
            const uint32_t q = newIndex[I.trParticle[t]];
            if (q >= p.size()) continue;
            id.push_back(q);
            tr.insert(tr.end(), I.trRest.begin() + 4 * t, I.trRest.begin() + 4 * t + 4);
            oc.push_back(I.trOcean[t]);
        }
        check(astraea_set_tracers(I.ctx, (uint32_t)id.size(), id.data(), tr.data(), oc.data()), "astraea_set_tracers");
    }
    check(astraea_relax(I.ctx, steps, nullptr, &nominal, &outFinalVrmsOverVesc), "astraea_relax");
    Particles r = downloadParticles(I.ctx, p.size());
    double rMax = 0.0;
    for (size_t i = 0; i < r.size(); ++i)
        rMax = std::max(rMax, std::sqrt((double)r.x[i] * r.x[i] + (double)r.y[i] * r.y[i] + (double)r.z[i] * r.z[i]));
    // tracers follow their particles into the merged order (the relaxed body first); those of
    // the relaxed body rest where it settled (simulation.cpp:1624-1648)
    for (uint32_t t = 0; t < I.tracerCount; ++t) {
        const uint32_t q = newIndex[I.trParticle[t]];
        I.trParticle[t] = q;
        if (q < p.size()) {
            const double ri = std::sqrt((double)r.x[q] * r.x[q] + (double)r.y[q] * r.y[q] + (double)r.z[q] * r.z[q]);
            I.trRest[4 * t] = (float)ri;
            if (ri > 1.0e-9) {
                I.trRest[4 * t + 1] = (float)(r.x[q] / ri);
                I.trRest[4 * t + 2] = (float)(r.y[q] / ri);
                I.trRest[4 * t + 3] = (float)(r.z[q] / ri);
            }
        }
    }
    if (!rest.empty()) {
        const double back[3] = {-c[0], -c[1], -c[2]}, backV[3] = {-v[0], -v[1], -v[2]};
        r.shift(back, backV);
    }
    if (rMax > 0.0) (bodyId == 0u ? I.targetRadiusPsu : I.impactorRadiusPsu) = rMax;
    if (!rest.empty()) r.append(all.subset(rest));
    uploadParticles(I.ctx, r, {(uint32_t)r.size()});
    I.startFromUpload((uint32_t)r.size());
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

// simulation.cpp:1805-1867: particle records about the system centre, tracer records about the
// living body's, both of the last telemetry.
void Simulation::extractRenderFrame(std::vector<uint8_t>& outParticleBytes, std::vector<uint8_t>& outTracerBytes,
                                    std::vector<uint8_t>* outSpeedBytes) {
    Impl& I = *impl_;
    if (I.count == 0u) {
        outParticleBytes.clear();
        outTracerBytes.clear();
        if (outSpeedBytes) outSpeedBytes->clear();
        return;
    }
    outParticleBytes.resize(20u * (size_t)I.count);
    if (outSpeedBytes) outSpeedBytes->resize(I.count);
    check(astraea_pack_render(I.ctx, I.sysCom, outParticleBytes.data(), outSpeedBytes ? outSpeedBytes->data() : nullptr),
          "astraea_pack_render");
    if (I.tracerCount > 0u) {
        outTracerBytes.resize(40u * (size_t)I.tracerCount);
        check(astraea_extract_tracers(I.ctx, I.livingCom, I.sysCom, outTracerBytes.data()), "astraea_extract_tracers");
    } else {
        outTracerBytes.clear();
    }
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
