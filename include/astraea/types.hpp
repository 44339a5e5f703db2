// types.hpp — the public value types of the reference's Simulation API (engine/types.hpp:10-111),
// kept identical (names, enumerator values, field order and defaults) for the drop-in.
// The device descriptor ParticleBuffersDevice (types.hpp:114-161) is not part of this API: the
// device layout lives behind the C ABI (include/astraea_hip.h).
#pragma once

#include <cstdint>
#include <string>

#include "units.hpp"

namespace astraea {

enum class MaterialId : uint8_t { Iron = 0, Granite = 1, WaterIce = 2, HydrogenHelium = 3, Regolith = 4 };

enum class PhaseId : uint8_t { Solid = 0, Molten = 1, Vapor = 2, Plasma = 3, LiquidWater = 4, Gas = 5 };

enum class PlanetArchetype : uint8_t {
    LivingTerra = 0, IronCore = 1, OceanIce = 2, LavaIo = 3, GasGiant = 4, ProtoTheia = 5
};

enum class ScenarioPresetId : uint8_t {
    EarthTheiaMoonForming = 0,
    MercuryMantleStrip = 1,
    OceanWorldSteamPlume = 2,
    GasGiantCoreImpact = 3,
    LivingGaiaAsteroidCataclysm = 4
};

enum class CollisionOutcome : uint8_t {
    Approaching = 0, Merging = 1, GrazeAndMerge = 2, HitAndRun = 3, CatastrophicDisruption = 4
};

struct SlicePlaneConfig {
    bool enabled = false;
    float normalX = 0.0f;
    float normalY = 1.0f;
    float normalZ = 0.0f;
    float offsetKm = 0.0f;
};

struct BodyInitSpec {
    PlanetArchetype archetype = PlanetArchetype::LivingTerra;
    double massKg = units::M_EARTH_SI;
    double surfaceTempK = 288.0;
    double coreTempK = 5500.0;
    double spinOmegaRadS = 7.292115e-5;
    uint32_t relaxSteps = 80;
};

struct CollisionScenarioConfig {
    std::string presetName = "LivingGaiaAsteroidCataclysm";
    BodyInitSpec target = {};
    BodyInitSpec impactor = {PlanetArchetype::ProtoTheia, 0.025 * units::M_EARTH_SI, 240.0, 1800.0, 0.0, 80};
    double impactVelocityRatioVesc = 1.65;
    double impactParamB = 0.30;
    double initialSepRadii = 3.2;
    uint32_t totalParticles = 100000;
    uint32_t tracerGridRes = 16;
    uint64_t randomSeed = 42;
};

struct TelemetrySnapshot {
    uint64_t frameSeq = 0;
    double simTimeSec = 0.0;
    float dtSec = 0.0f;
    CollisionOutcome outcome = CollisionOutcome::Approaching;
    double kineticEnergyJ = 0.0;
    double potentialEnergyJ = 0.0;
    double internalEnergyJ = 0.0;
    double radiatedEnergyJ = 0.0;
    double totalEnergyJ = 0.0;
    double initialTotalEnergyJ = 0.0;
    double energyDriftRel = 0.0;
    double momentumKgMS[3] = {0.0, 0.0, 0.0};
    double angularMomentumKgM2S[3] = {0.0, 0.0, 0.0};
    double virialRatio = 0.0;
    double maxTempK = 0.0;
    double maxPressureGPa = 0.0;
    double boundMassFrac = 1.0;
    double diskMassFrac = 0.0;
    double ejectaMassFrac = 0.0;
    double qRatioQrd = 0.0;
    double gpuStepMs = 0.0;
    double livingBodyComKm[3] = {0.0, 0.0, 0.0};
    double livingBodyRadiusKm = 0.0;
    double livingBodySpinRadS = 0.0;
    double impactorComKm[3] = {0.0, 0.0, 0.0};
    uint32_t particleCount = 0;
    uint32_t tracerCount = 0;
    uint32_t tracerGridRes = 0;
    bool livingBodyPresent = false;
};

}  // namespace astraea
