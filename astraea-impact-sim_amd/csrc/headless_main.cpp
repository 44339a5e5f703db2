// headless_main.cpp — sim_headless: the reference's command-line runner (engine/headless_main.cpp:98-288)
// on the gfx950 Simulation.  Same flags, summary line and --verify-conservation gates
// (|dE|/|E0| < 0.01, |dP|/(M v_esc) < 1e-4, |dL|/|L0| < 0.005; exit 1 on failure or non-finite
// telemetry, 2 on an exception).  The mutual escape speed uses the radii of the bodies'
// hydrostatic profiles, as the reference's does.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/astraea/presets.hpp"
#include "../../include/astraea/simulation.hpp"
#include "../../include/astraea_hip.h"

using namespace astraea;

namespace {

// Radius of the body's hydrostatic profile, metres (headless_main.cpp:169-177 solves the same).
double bodyRadius(const BodyInitSpec& b) {
    const ArchetypeComposition& comp = getArchetypeComposition(b.archetype);
    astraea_planet_spec ps;
    astraea_default_planet_spec(&ps);
    ps.mass_kg = b.massKg;
    ps.num_layers = 0;
    for (uint32_t k = 0; k < comp.layerCount; ++k) {
        if (comp.layers[k].massFraction <= 0.0) continue;
        ps.layer_material[ps.num_layers] = (uint8_t)comp.layers[k].material;
        ps.layer_mass_frac[ps.num_layers] = comp.layers[k].massFraction;
        ++ps.num_layers;
    }
    ps.surface_temp_k = b.surfaceTempK;
    ps.core_temp_k = b.coreTempK;
    ps.surface_pressure_pa = b.archetype == PlanetArchetype::GasGiant ? 1.0e5 : 0.0;
    astraea_planet_info info{};
    if (astraea_planet_profile(&ps, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, &info) != ASTRAEA_OK)
        throw std::invalid_argument(std::string("hydrostatic profile failed: ") + astraea_last_error());
    return info.radius_m;
}

double norm3(const double v[3]) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

bool finiteSnapshot(const TelemetrySnapshot& s) {
    const double v[] = {s.simTimeSec, (double)s.dtSec, s.kineticEnergyJ, s.potentialEnergyJ, s.internalEnergyJ,
                        s.radiatedEnergyJ, s.totalEnergyJ, s.initialTotalEnergyJ, s.energyDriftRel,
                        s.momentumKgMS[0], s.momentumKgMS[1], s.momentumKgMS[2], s.angularMomentumKgM2S[0],
                        s.angularMomentumKgM2S[1], s.angularMomentumKgM2S[2], s.virialRatio, s.maxTempK,
                        s.maxPressureGPa, s.boundMassFrac, s.diskMassFrac, s.ejectaMassFrac, s.qRatioQrd, s.gpuStepMs,
                        s.livingBodyComKm[0], s.livingBodyComKm[1], s.livingBodyComKm[2], s.livingBodyRadiusKm,
                        s.livingBodySpinRadS, s.impactorComKm[0], s.impactorComKm[1], s.impactorComKm[2]};
    for (double d : v)
        if (!std::isfinite(d)) return false;
    return true;
}

const char* outcomeName(CollisionOutcome o) {
    switch (o) {
        case CollisionOutcome::Merging: return "Merging";
        case CollisionOutcome::GrazeAndMerge: return "GrazeAndMerge";
        case CollisionOutcome::HitAndRun: return "HitAndRun";
        case CollisionOutcome::CatastrophicDisruption: return "CatastrophicDisruption";
        default: return "Approaching";
    }
}

void usage(const char* argv0) {
    std::cout << "Usage: " << argv0
              << " [--preset NAME] [--particles N] [--steps N] [--dt-max SEC] [--relax-steps N]\n"
                 "       [--dump-telemetry FILE.json] [--output-csv FILE.csv] [--verify-conservation]"
                 " [--list-presets]\n";
}

bool writeFile(const std::string& path, const std::string& text) {
    std::ofstream f(path, std::ios::binary);
    if (!f) return false;
    f << text;
    return (bool)f;
}

}  // namespace

int main(int argc, char** argv) {
    std::string presetName = "LivingGaiaAsteroidCataclysm";
    uint32_t particles = 100000, steps = 100;
    double dtMaxSec = 5.0;
    bool relaxOverride = false, verify = false;
    uint32_t relaxSteps = 80;
    std::string jsonPath, csvPath;
    auto need = [&](int& i, const char* flag) -> std::string {
        if (i + 1 >= argc) throw std::invalid_argument(std::string("Missing value for ") + flag);
        return std::string(argv[++i]);
    };
    try {
        for (int i = 1; i < argc; ++i) {
            const std::string a = argv[i];
            if (a == "--help" || a == "-h") { usage(argv[0]); return 0; }
            if (a == "--list-presets") {
                for (int k = 0; k <= (int)ScenarioPresetId::LivingGaiaAsteroidCataclysm; ++k)
                    std::cout << presetToString((ScenarioPresetId)k) << '\n';
                return 0;
            }
            if (a == "--preset") presetName = need(i, "--preset");
            else if (a == "--particles") {
                const long long v = std::stoll(need(i, "--particles"));
                if (v <= 0) throw std::invalid_argument("--particles must be > 0");
                particles = (uint32_t)v;
            } else if (a == "--steps") {
                const long long v = std::stoll(need(i, "--steps"));
                if (v < 0) throw std::invalid_argument("--steps must be >= 0");
                steps = (uint32_t)v;
            } else if (a == "--dt-max") {
                dtMaxSec = std::stod(need(i, "--dt-max"));
                if (!(dtMaxSec > 0.0) || !std::isfinite(dtMaxSec)) throw std::invalid_argument("--dt-max must be > 0");
            } else if (a == "--relax-steps") {
                const long long v = std::stoll(need(i, "--relax-steps"));
                if (v < 0) throw std::invalid_argument("--relax-steps must be >= 0");
                relaxOverride = true;
                relaxSteps = (uint32_t)v;
            } else if (a == "--dump-telemetry") jsonPath = need(i, "--dump-telemetry");
            else if (a == "--output-csv") csvPath = need(i, "--output-csv");
            else if (a == "--verify-conservation") verify = true;
            else throw std::invalid_argument("Unknown option: " + a);
        }
        CollisionScenarioConfig cfg = makePresetConfig(parsePresetName(presetName), particles);
        if (relaxOverride) cfg.target.relaxSteps = cfg.impactor.relaxSteps = relaxSteps;
        const double mTot = cfg.target.massKg + std::max(0.0, cfg.impactor.massKg);
        const double rScale = std::max(bodyRadius(cfg.target) + (cfg.impactor.massKg > 0 ? bodyRadius(cfg.impactor) : 0.0), 1.0);
        const double vEsc = std::sqrt(2.0 * units::G_SI * mTot / rScale);

        Simulation sim(particles);
        sim.loadScenario(cfg);
        const TelemetrySnapshot t0 = sim.computeTelemetry();
        if (!finiteSnapshot(t0)) { std::cerr << "ERROR: Non-finite initial telemetry at t=0\n"; return 1; }
        std::vector<double> ms;
        TelemetrySnapshot last = t0;
        for (uint32_t s = 0; s < steps; ++s) {
            last = sim.step(dtMaxSec);
            ms.push_back(last.gpuStepMs);
            if (!finiteSnapshot(last)) { std::cerr << "ERROR: Non-finite telemetry at step " << (s + 1) << '\n'; return 1; }
        }
        double med = 0.0;
        if (!ms.empty()) {
            std::sort(ms.begin(), ms.end());
            const size_t k = ms.size() / 2;
            med = ms.size() % 2 == 0 ? 0.5 * (ms[k - 1] + ms[k]) : ms[k];
        }
        const double dP[3] = {last.momentumKgMS[0] - t0.momentumKgMS[0], last.momentumKgMS[1] - t0.momentumKgMS[1],
                              last.momentumKgMS[2] - t0.momentumKgMS[2]};
        const double dL[3] = {last.angularMomentumKgM2S[0] - t0.angularMomentumKgM2S[0],
                              last.angularMomentumKgM2S[1] - t0.angularMomentumKgM2S[1],
                              last.angularMomentumKgM2S[2] - t0.angularMomentumKgM2S[2]};
        const double pRel = mTot * vEsc > 0.0 ? norm3(dP) / (mTot * vEsc) : 0.0;
        const double lRel = norm3(dL) / std::max(norm3(t0.angularMomentumKgM2S), mTot * rScale * vEsc * 1.0e-6);
        if (!jsonPath.empty() && !writeFile(jsonPath, sim.exportTelemetryJson())) {
            std::cerr << "ERROR: Failed to write JSON telemetry to " << jsonPath << '\n';
            return 1;
        }
        if (!csvPath.empty() && !writeFile(csvPath, sim.exportTelemetryCsv())) {
            std::cerr << "ERROR: Failed to write CSV telemetry to " << csvPath << '\n';
            return 1;
        }
        std::cout << std::setprecision(6) << "preset=" << cfg.presetName << " particles=" << last.particleCount
                  << " tracers=" << last.tracerCount << " steps=" << steps << " simTimeSec=" << last.simTimeSec
                  << " dtSec=" << last.dtSec << " medianGpuStepMs=" << med << " energyDriftRel=" << last.energyDriftRel
                  << " dPRel=" << pRel << " dLRel=" << lRel << " maxTempK=" << last.maxTempK
                  << " maxPressureGPa=" << last.maxPressureGPa << " boundMassFrac=" << last.boundMassFrac
                  << " diskMassFrac=" << last.diskMassFrac << " ejectaMassFrac=" << last.ejectaMassFrac
                  << " qRatioQrd=" << last.qRatioQrd << " outcome=" << outcomeName(last.outcome) << '\n';
        if (verify) {
            bool ok = true;
            if (!(std::abs(last.energyDriftRel) < 0.01)) {
                std::cerr << "CONSERVATION FAIL: |energyDriftRel| = " << std::abs(last.energyDriftRel) << " >= 0.01\n";
                ok = false;
            }
            if (!(pRel < 1.0e-4)) {
                std::cerr << "CONSERVATION FAIL: |dP|/(M*v_esc) = " << pRel << " >= 1e-4\n";
                ok = false;
            }
            if (!(lRel < 0.005)) {
                std::cerr << "CONSERVATION FAIL: |dL|/|L0| = " << lRel << " >= 0.005\n";
                ok = false;
            }
            if (!ok) return 1;
        }
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "ERROR: " << e.what() << '\n';
        return 2;
    }
}
