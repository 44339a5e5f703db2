// test_simulation.cpp — acceptance cases of the reference's C++ tests run end to end through
// astraea::Simulation on the GPU, same configurations and expectations:
//   outcomes  tests/cpp/test_conservation.cpp:479-541 (AC21): collision-outcome classification;
//   relax     tests/cpp/test_planet_builder.cpp:329-360: relaxBody(0, 80) leaves every archetype
//             with v_rms / v_esc < 0.01.
// Usage: test_simulation {outcomes|relax}.  Prints one line per case; the exit status is the
// number of failed expectations (200 + k on an exception in case k).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>

#include "../../include/astraea/presets.hpp"
#include "../../include/astraea/simulation.hpp"

using namespace astraea;

namespace {

const char* name(CollisionOutcome o) {
    switch (o) {
        case CollisionOutcome::Approaching: return "Approaching";
        case CollisionOutcome::Merging: return "Merging";
        case CollisionOutcome::GrazeAndMerge: return "GrazeAndMerge";
        case CollisionOutcome::HitAndRun: return "HitAndRun";
        case CollisionOutcome::CatastrophicDisruption: return "CatastrophicDisruption";
    }
    return "?";
}

int fails = 0;

void report(const char* label, const TelemetrySnapshot& t, CollisionOutcome want) {
    const bool ok = t.outcome == want;
    fails += ok ? 0 : 1;
    std::printf("%s outcome=%s expected=%s bound=%.4f disk=%.4f ejecta=%.4f t=%.1f s frames=%llu %s\n", label,
                name(t.outcome), name(want), t.boundMassFrac, t.diskMassFrac, t.ejectaMassFrac, t.simTimeSec,
                (unsigned long long)t.frameSeq, ok ? "ok" : "FAIL");
    std::fflush(stdout);
}

void expect(bool ok, const char* what) {
    if (!ok) {
        ++fails;
        std::printf("  FAIL %s\n", what);
        std::fflush(stdout);
    }
}

}  // namespace

int outcomes() {
    int stage = 0;
    try {
        Simulation sim(8000);
        {  // 0) pre-impact state classifies as Approaching
            sim.loadScenario(makePresetConfig(ScenarioPresetId::EarthTheiaMoonForming, 4000));
            report("approaching", sim.computeTelemetry(), CollisionOutcome::Approaching);
        }
        stage = 1;
        {  // 1) low-energy head-on / gentle impact -> Merging
            CollisionScenarioConfig cfg = makePresetConfig(ScenarioPresetId::OceanWorldSteamPlume, 5000);
            cfg.impactVelocityRatioVesc = 0.95;
            cfg.impactParamB = 0.05;
            cfg.initialSepRadii = 1.25;
            sim.loadScenario(cfg);
            sim.stepForDuration(1600.0, 10.0);
            const TelemetrySnapshot t = sim.computeTelemetry();
            report("merging", t, CollisionOutcome::Merging);
            expect(t.boundMassFrac > 0.70, "merging: boundMassFrac > 0.70");
        }
        stage = 2;
        {  // 2) oblique fast grazing impact -> HitAndRun
            CollisionScenarioConfig cfg;
            cfg.presetName = "HitAndRunCheck";
            cfg.target = {PlanetArchetype::ProtoTheia, 0.50 * units::M_EARTH_SI, 240.0, 2200.0, 0.0, 60};
            cfg.impactor = {PlanetArchetype::ProtoTheia, 0.35 * units::M_EARTH_SI, 240.0, 2000.0, 0.0, 60};
            cfg.impactVelocityRatioVesc = 1.85;
            cfg.impactParamB = 0.80;
            cfg.initialSepRadii = 1.40;
            cfg.totalParticles = 5000;
            cfg.tracerGridRes = 8;
            cfg.randomSeed = 42;
            sim.loadScenario(cfg);
            sim.stepForDuration(1600.0, 10.0);
            report("hit_and_run", sim.computeTelemetry(), CollisionOutcome::HitAndRun);
        }
        stage = 3;
        {  // 3) high-energy head-on shattering collision -> CatastrophicDisruption
            CollisionScenarioConfig cfg;
            cfg.presetName = "CatastrophicDisruptionCheck";
            cfg.target = {PlanetArchetype::ProtoTheia, 0.45 * units::M_EARTH_SI, 440.0, 2200.0, 0.0, 60};
            cfg.impactor = {PlanetArchetype::ProtoTheia, 0.45 * units::M_EARTH_SI, 440.0, 2200.0, 0.0, 60};
            cfg.impactVelocityRatioVesc = 4.20;
            cfg.impactParamB = 0.05;
            cfg.initialSepRadii = 1.25;
            cfg.totalParticles = 5000;
            cfg.tracerGridRes = 8;
            cfg.randomSeed = 42;
            sim.loadScenario(cfg);
            sim.stepForDuration(1500.0, 10.0);
            const TelemetrySnapshot t = sim.computeTelemetry();
            report("catastrophic", t, CollisionOutcome::CatastrophicDisruption);
            expect(t.boundMassFrac < 0.55, "catastrophic: boundMassFrac < 0.55");
        }
    } catch (const std::exception& e) {
        std::printf("exception in case %d: %s\n", stage, e.what());
        return 200 + stage;
    }
    return fails;
}

int relax() {
    const PlanetArchetype archetypes[6] = {PlanetArchetype::LivingTerra, PlanetArchetype::IronCore,
                                           PlanetArchetype::OceanIce, PlanetArchetype::LavaIo,
                                           PlanetArchetype::GasGiant, PlanetArchetype::ProtoTheia};
    int k = 0;
    try {
        Simulation sim(12000);
        for (; k < 6; ++k) {
            CollisionScenarioConfig cfg;
            cfg.presetName = "RelaxCheck";
            cfg.target = makeDefaultBodySpec(archetypes[k]);
            cfg.target.relaxSteps = 0;
            cfg.impactor = makeDefaultBodySpec(PlanetArchetype::ProtoTheia);
            cfg.impactor.massKg = 0.25 * cfg.target.massKg;
            cfg.impactor.relaxSteps = 0;
            cfg.impactVelocityRatioVesc = 1.0;
            cfg.impactParamB = 0.0;
            cfg.initialSepRadii = 10.0;
            cfg.totalParticles = 8000;
            cfg.tracerGridRes = 8;
            cfg.randomSeed = 42;
            sim.loadScenario(cfg);
            double vr = 1.0;
            sim.relaxBody(0, 80, vr);
            const bool ok = std::isfinite(vr) && vr < 0.01;
            fails += ok ? 0 : 1;
            std::printf("relax archetype=%d vrms_over_vesc=%.3e %s\n", k, vr, ok ? "ok" : "FAIL");
            std::fflush(stdout);
        }
    } catch (const std::exception& e) {
        std::printf("exception in archetype %d: %s\n", k, e.what());
        return 200 + k;
    }
    return fails;
}

int main(int argc, char** argv) {
    if (argc == 2 && std::strcmp(argv[1], "outcomes") == 0) return outcomes();
    if (argc == 2 && std::strcmp(argv[1], "relax") == 0) return relax();
    std::fprintf(stderr, "usage: %s {outcomes|relax}\n", argv[0]);
    return 2;
}
