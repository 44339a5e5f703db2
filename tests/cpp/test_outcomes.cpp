// test_outcomes.cpp — collision-outcome classification end to end through astraea::Simulation:
// the four cases of the reference's tests/cpp/test_conservation.cpp:479-541 (AC21), same
// configurations, durations and expectations.  Prints one line per case; the exit status is the
// number of failed expectations (2 x 100 + k on an exception in case k).
#include <cstdio>
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

int main() {
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
