// test_simulation.cpp — acceptance cases of the reference's C++ tests run end to end through
// astraea::Simulation on the GPU, same configurations and expectations:
//   outcomes  tests/cpp/test_conservation.cpp:479-541 (AC21): collision-outcome classification;
//   relax     tests/cpp/test_planet_builder.cpp:329-360: relaxBody(0, 80) leaves every archetype
//             with v_rms / v_esc < 0.01;
//   render    tests/cpp/test_conservation.cpp:260-369 (AC12): render-frame layout, quantisation and
//             the optional speed array, and the PCLS frame of it (server/binary_protocol.cpp);
//   pcls      the PCLS header and payload layout on synthetic records (no GPU needed).
// Usage: test_simulation {outcomes|relax|render|pcls}.  Prints one line per case; the exit status
// is the number of failed expectations (200 + k on an exception in case k).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>

#include <vector>

#include "../../include/astraea/binary_protocol.hpp"
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

float f32(const uint8_t* p) {
    float v;
    std::memcpy(&v, p, 4);
    return v;
}
uint16_t u16(const uint8_t* p) { return (uint16_t)(p[0] | (p[1] << 8)); }
uint32_t u32(const uint8_t* p) { return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24); }

// The frame of (particles, tracers, speed) decodes to the same header fields and payload.
void checkFrame(const TelemetrySnapshot& snap, const std::vector<uint8_t>& p, const std::vector<uint8_t>& t,
                const std::vector<uint8_t>* speed, bool zip) {
    using namespace astraea::server;
    const std::vector<uint8_t> frame = buildBinaryFrame(snap, p, t, speed, zip);
    BinaryFrameHeader h;
    expect(decodeHeader40(frame.data(), frame.size(), h), "frame header decodes");
    const bool withSpeed = speed && !p.empty() && speed->size() == p.size() / 20;
    std::vector<uint8_t> raw(p);
    raw.insert(raw.end(), t.begin(), t.end());
    if (withSpeed) raw.insert(raw.end(), speed->begin(), speed->end());
    expect(h.particleCount == p.size() / 20 && h.tracerCount == t.size() / 40, "frame counts");
    expect(h.payloadBytes == raw.size(), "payloadBytes is the raw payload size");
    expect(((h.flags & kFlagHasSpeedArray) != 0) == withSpeed, "speed flag");
    expect(((h.flags & kFlagHasTracers) != 0) == !t.empty(), "tracer flag");
    expect(((h.flags & kFlagHasLivingCom) != 0) == snap.livingBodyPresent, "living flag");
    expect(h.tracerGridRes == (t.empty() ? 0u : snap.tracerGridRes), "tracer grid");
    expect(h.simTimeSec == snap.simTimeSec && h.dtSec == snap.dtSec && h.frameSeq == (uint32_t)snap.frameSeq,
           "frame time fields");
    std::vector<uint8_t> body(frame.begin() + kPclsHeaderBytes, frame.end());
    if (h.flags & kFlagZlib) {
        std::vector<uint8_t> inflated;
        expect(zip && zlibDecompress(body.data(), body.size(), h.payloadBytes, inflated), "zlib payload inflates");
        body = inflated;
    } else {
        expect(!zip, "a zlib frame is flagged");
    }
    expect(body == raw, "payload = particles | tracers | speed");
}

int render() {
    int stage = 0;
    try {
        const CollisionScenarioConfig cfg = makePresetConfig(ScenarioPresetId::LivingGaiaAsteroidCataclysm, 100000);
        Simulation sim(105000);
        sim.loadScenario(cfg);
        sim.step(5.0);
        sim.step(5.0);
        const TelemetrySnapshot snap = sim.computeTelemetry();
        expect(snap.particleCount == 100000u, "particleCount 100000");
        expect(snap.tracerCount == 1536u, "tracerCount 6 x 16^2");
        expect(snap.livingBodyPresent, "living body present");
        std::vector<uint8_t> p0, t0, p1, t1, sp;
        sim.extractRenderFrame(p0, t0, nullptr);
        sim.extractRenderFrame(p1, t1, &sp);
        expect(p0.size() == 20u * 100000u && t0.size() == 40u * 1536u, "record sizes");
        expect(sp.size() == 100000u, "speed size");
        expect(p0 == p1 && t0 == t1, "speed bytes change nothing else");
        const double ln20000 = std::log(20000.0), ln1e7 = std::log(1.0e7), ln2000 = std::log(2000.0);
        double sx = 0, sy = 0, sz = 0, maxT = 0, maxP = 0, maxImpSpeed = 0;
        uint32_t b0 = 0, b1 = 0;
        bool fieldsOk = true;
        for (size_t i = 0; i < 100000u; ++i) {
            const uint8_t* r = p1.data() + 20 * i;
            const float x = f32(r), y = f32(r + 4), z = f32(r + 8);
            fieldsOk = fieldsOk && std::isfinite(x) && std::isfinite(y) && std::isfinite(z);
            sx += x; sy += y; sz += z;
            const uint8_t body = r[17] >> 4, phase = r[17] & 0x0f;
            fieldsOk = fieldsOk && r[16] <= 4 && phase <= 5 && body <= 1;
            (body == 0 ? b0 : b1) += 1;
            maxT = std::max(maxT, 10.0 * std::exp(u16(r + 12) / 65535.0 * ln20000));
            const uint16_t pc = u16(r + 14);
            maxP = std::max(maxP, pc == 0 ? 0.0 : 1.0e-3 * std::exp((pc - 1) / 65534.0 * ln1e7));
            const double rad = std::exp(r[19] / 255.0 * ln2000);
            fieldsOk = fieldsOk && rad >= 1.0 && rad <= 2000.0 + 1e-6;
            const double spd = 0.05 * std::exp(sp[i] / 255.0 * ln2000);
            if (body == 1) maxImpSpeed = std::max(maxImpSpeed, spd);
        }
        expect(fieldsOk, "record fields in range");
        expect(b0 > 90000u && b1 > 1000u, "body counts");
        expect(maxImpSpeed > 5.0, "impactor speed > 5 km/s");
        const double meanKm = std::sqrt(sx * sx + sy * sy + sz * sz) / 100000.0;
        expect(meanKm < 500.0, "records centred on the system COM");
        if (snap.maxTempK >= 10.0 && snap.maxTempK <= 200000.0)
            expect(std::fabs(maxT - snap.maxTempK) <= 1e-3 * snap.maxTempK, "temperature quantisation < 0.1 %");
        if (snap.maxPressureGPa >= 1.0e-3 && snap.maxPressureGPa <= 1.0e4)
            expect(std::fabs(maxP - snap.maxPressureGPa) <= 1e-3 * snap.maxPressureGPa, "pressure quantisation < 0.1 %");
        std::vector<uint32_t> seen(1536, 0);
        for (size_t k = 0; k < 1536; ++k) {
            const uint16_t idx = u16(t1.data() + 40 * k);
            if (idx < 1536) seen[idx] += 1;
        }
        bool unique = true;
        for (uint32_t c : seen) unique = unique && c == 1;
        expect(unique, "tracer records carry every index once");
        std::printf("render maxT=%.1f (snap %.1f) maxP=%.4g (snap %.4g) maxImpactorSpeed=%.2f km/s mean=%.2f km\n",
                    maxT, snap.maxTempK, maxP, snap.maxPressureGPa, maxImpSpeed, meanKm);
        stage = 1;
        checkFrame(snap, p1, t1, &sp, false);
        checkFrame(snap, p1, t1, &sp, true);
        checkFrame(snap, p0, t0, nullptr, true);
        stage = 2;
        Simulation other(5000);
        other.loadScenario(makePresetConfig(ScenarioPresetId::MercuryMantleStrip, 4000));
        std::vector<uint8_t> p2, t2;
        other.extractRenderFrame(p2, t2, nullptr);
        expect(p2.size() == 20u * 4000u && t2.empty(), "no tracers without a living body");
    } catch (const std::exception& e) {
        std::printf("exception in render stage %d: %s\n", stage, e.what());
        return 200 + stage;
    }
    std::printf("render %s\n", fails ? "FAIL" : "ok");
    return fails;
}

// Header bytes at their offsets (the layout the web client decodes, web/tests/binary_protocol.test.ts).
int pcls() {
    using namespace astraea::server;
    TelemetrySnapshot snap;
    snap.simTimeSec = 12.5;
    snap.frameSeq = 42;
    snap.dtSec = 0.25f;
    snap.tracerGridRes = 16;
    snap.livingBodyPresent = true;
    std::vector<uint8_t> p(3 * 20), t(2 * 40), sp(3);
    for (size_t i = 0; i < p.size(); ++i) p[i] = (uint8_t)(i * 7 + 1);
    for (size_t i = 0; i < t.size(); ++i) t[i] = (uint8_t)(255 - i);
    sp = {9, 8, 7};
    const std::vector<uint8_t> f = buildBinaryFrame(snap, p, t, &sp, false);
    expect(f.size() == 40 + 60 + 80 + 3, "frame size");
    expect(f[0] == 0x53 && f[1] == 0x4C && f[2] == 0x43 && f[3] == 0x50, "magic bytes");
    expect(u16(f.data() + 4) == 1, "version");
    expect(u16(f.data() + 6) == (kFlagHasTracers | kFlagHasLivingCom | kFlagHasSpeedArray), "flags");
    double tsec;
    std::memcpy(&tsec, f.data() + 8, 8);
    expect(tsec == 12.5 && u32(f.data() + 16) == 42 && f32(f.data() + 20) == 0.25f, "time fields");
    expect(u32(f.data() + 24) == 3 && u32(f.data() + 28) == 2 && u32(f.data() + 32) == 16 && u32(f.data() + 36) == 143,
           "counts and payload size");
    expect(std::equal(p.begin(), p.end(), f.begin() + 40) && std::equal(t.begin(), t.end(), f.begin() + 100) &&
               std::equal(sp.begin(), sp.end(), f.begin() + 180),
           "payload order");
    BinaryFrameHeader h;
    std::vector<uint8_t> bad(f);
    bad[0] ^= 1;
    expect(!decodeHeader40(bad.data(), bad.size(), h), "wrong magic rejected");
    bad = f;
    bad[4] = 2;
    expect(!decodeHeader40(bad.data(), bad.size(), h), "wrong version rejected");
    expect(!decodeHeader40(f.data(), 39, h), "short header rejected");
    std::vector<uint8_t> spBad = {1, 2};
    const std::vector<uint8_t> g = buildBinaryFrame(snap, p, t, &spBad, false);
    expect(u32(g.data() + 36) == 140 && (u16(g.data() + 6) & kFlagHasSpeedArray) == 0, "speed of the wrong size dropped");
    checkFrame(snap, p, t, &sp, true);
    std::vector<uint8_t> z, back;
    expect(zlibCompress(nullptr, 0, z) && zlibDecompress(z.data(), z.size(), 0, back) && back.empty(), "empty zlib");
    expect(!zlibDecompress(z.data(), z.size(), 5, back), "size mismatch rejected");
    std::printf("pcls %s\n", fails ? "FAIL" : "ok");
    return fails;
}

int main(int argc, char** argv) {
    if (argc == 2 && std::strcmp(argv[1], "outcomes") == 0) return outcomes();
    if (argc == 2 && std::strcmp(argv[1], "relax") == 0) return relax();
    if (argc == 2 && std::strcmp(argv[1], "render") == 0) return render();
    if (argc == 2 && std::strcmp(argv[1], "pcls") == 0) return pcls();
    std::fprintf(stderr, "usage: %s {outcomes|relax|render|pcls}\n", argv[0]);
    return 2;
}
