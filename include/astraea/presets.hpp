// presets.hpp — archetype compositions and scenario presets of the reference's public API
// (engine/presets.hpp:1-33), same declarations so callers compile unchanged.
#pragma once

#include <array>
#include <cstdint>
#include <string>

#include "types.hpp"

namespace astraea {

struct PlanetLayerDef {
    MaterialId material = MaterialId::Granite;
    double massFraction = 0.0;
};

struct ArchetypeComposition {
    std::array<PlanetLayerDef, 3> layers{};   // innermost first
    uint32_t layerCount = 0;
    double defaultMassKg = units::M_EARTH_SI;
    double defaultSurfaceTempK = 288.0;
    double defaultCoreTempK = 5500.0;
    double defaultSpinOmegaRadS = 0.0;
};

// Out-of-range archetypes map to LivingTerra (presets.cpp:92-98).
const ArchetypeComposition& getArchetypeComposition(PlanetArchetype archetype);
BodyInitSpec makeDefaultBodySpec(PlanetArchetype archetype);

CollisionScenarioConfig makePresetConfig(ScenarioPresetId preset, uint32_t totalParticles = 100000);
ScenarioPresetId parsePresetName(const std::string& name);  // throws std::invalid_argument on unknown
std::string presetToString(ScenarioPresetId id);

}  // namespace astraea
