// presets.cpp — archetype table and scenario presets (include/astraea/presets.hpp).
//
// The numbers are the reference's (engine/presets.cpp:9-88 archetypes, 147-210 presets): they
// define the scenarios, so they are kept as they are.
#include <stdexcept>

#include "../../include/astraea/presets.hpp"

namespace astraea {
namespace {

using M = MaterialId;

ArchetypeComposition comp(std::array<PlanetLayerDef, 3> layers, uint32_t count, double massKg, double surfaceK,
                          double coreK, double spin) {
    ArchetypeComposition c;
    c.layers = layers;
    c.layerCount = count;
    c.defaultMassKg = massKg;
    c.defaultSurfaceTempK = surfaceK;
    c.defaultCoreTempK = coreK;
    c.defaultSpinOmegaRadS = spin;
    return c;
}

const ArchetypeComposition kArchetypes[6] = {
    comp({{{M::Iron, 0.320}, {M::Granite, 0.675}, {M::WaterIce, 0.005}}}, 3, 5.97219e24, 288.0, 5500.0, 7.292115e-5),
    comp({{{M::Iron, 0.680}, {M::Granite, 0.320}, {M::Granite, 0.0}}}, 2, 3.3011e23, 440.0, 3200.0, 1.24e-6),
    comp({{{M::Iron, 0.150}, {M::Granite, 0.450}, {M::WaterIce, 0.400}}}, 3, 4.0e24, 275.0, 3800.0, 5.0e-5),
    comp({{{M::Iron, 0.250}, {M::Granite, 0.750}, {M::Granite, 0.0}}}, 2, 2.0e24, 1550.0, 4200.0, 4.11e-5),
    comp({{{M::Iron, 0.050}, {M::Granite, 0.100}, {M::HydrogenHelium, 0.850}}}, 3, 1.2e26, 165.0, 18000.0, 1.76e-4),
    comp({{{M::Iron, 0.220}, {M::Granite, 0.580}, {M::Regolith, 0.200}}}, 3, 6.417e23, 240.0, 2200.0, 7.09e-5),
};

struct PresetDef {
    ScenarioPresetId id;
    const char* name;
    PlanetArchetype target, impactor;
    double targetEarths, impactorEarths, vRatio, b, sep;
};

const PresetDef kPresets[5] = {
    {ScenarioPresetId::EarthTheiaMoonForming, "EarthTheiaMoonForming", PlanetArchetype::LivingTerra,
     PlanetArchetype::ProtoTheia, 1.0, 0.13, 1.05, 0.72, 3.0},
    {ScenarioPresetId::MercuryMantleStrip, "MercuryMantleStrip", PlanetArchetype::ProtoTheia,
     PlanetArchetype::IronCore, 0.45, 0.22, 2.80, 0.35, 3.2},
    {ScenarioPresetId::OceanWorldSteamPlume, "OceanWorldSteamPlume", PlanetArchetype::OceanIce,
     PlanetArchetype::ProtoTheia, 0.75, 0.15, 1.35, 0.25, 3.2},
    {ScenarioPresetId::GasGiantCoreImpact, "GasGiantCoreImpact", PlanetArchetype::GasGiant, PlanetArchetype::IronCore,
     20.0, 1.0, 1.40, 0.15, 2.8},
    {ScenarioPresetId::LivingGaiaAsteroidCataclysm, "LivingGaiaAsteroidCataclysm", PlanetArchetype::LivingTerra,
     PlanetArchetype::ProtoTheia, 1.0, 0.025, 1.65, 0.30, 3.2},
};

}  // namespace

const ArchetypeComposition& getArchetypeComposition(PlanetArchetype archetype) {
    const uint8_t k = static_cast<uint8_t>(archetype);
    return kArchetypes[k > 5 ? 0 : k];
}

BodyInitSpec makeDefaultBodySpec(PlanetArchetype archetype) {
    const ArchetypeComposition& c = getArchetypeComposition(archetype);
    BodyInitSpec s;
    s.archetype = archetype;
    s.massKg = c.defaultMassKg;
    s.surfaceTempK = c.defaultSurfaceTempK;
    s.coreTempK = c.defaultCoreTempK;
    s.spinOmegaRadS = c.defaultSpinOmegaRadS;
    s.relaxSteps = 80;
    return s;
}

std::string presetToString(ScenarioPresetId id) {
    for (const PresetDef& p : kPresets)
        if (p.id == id) return p.name;
    throw std::invalid_argument("Unknown ScenarioPresetId");
}

ScenarioPresetId parsePresetName(const std::string& name) {
    for (const PresetDef& p : kPresets)
        if (name == p.name) return p.id;
    throw std::invalid_argument("Unknown preset name: " + name);
}

CollisionScenarioConfig makePresetConfig(ScenarioPresetId preset, uint32_t totalParticles) {
    const PresetDef* def = nullptr;
    for (const PresetDef& p : kPresets)
        if (p.id == preset) def = &p;
    if (!def) throw std::invalid_argument("Unknown ScenarioPresetId");
    CollisionScenarioConfig c;
    c.presetName = def->name;
    c.totalParticles = totalParticles;
    c.tracerGridRes = 16;
    c.randomSeed = 42;
    c.target = makeDefaultBodySpec(def->target);
    c.target.massKg = def->targetEarths * units::M_EARTH_SI;
    c.impactor = makeDefaultBodySpec(def->impactor);
    c.impactor.massKg = def->impactorEarths * units::M_EARTH_SI;
    c.impactVelocityRatioVesc = def->vRatio;
    c.impactParamB = def->b;
    c.initialSepRadii = def->sep;
    // per-preset temperature and spin overrides (presets.cpp:168, 201-202)
    if (preset == ScenarioPresetId::MercuryMantleStrip) c.target.surfaceTempK = 440.0;
    if (preset == ScenarioPresetId::LivingGaiaAsteroidCataclysm) {
        c.impactor.coreTempK = 1800.0;
        c.impactor.spinOmegaRadS = 0.0;
    }
    return c;
}

}  // namespace astraea
