// ref_eos.cpp — C entry points over the reference's own host code: the EOS (engine/eos.cu,
// eos.hpp) and the hydrostatic profile solver (engine/planet_builder.cpp with presets.cpp), which
// oracle/Makefile compiles as C++ from the reference tree into oracle/_ref/libref_eos.so.
//
// TEST INFRASTRUCTURE ONLY: tests/test_oracle_ref_eos.py checks the oracle's EOS restatement
// and the library's planet builder against it.  Nothing here is copied from the reference; these
// wrappers only convert the reference's C++ signatures (eos.hpp:47-64, planet_builder.hpp:35-48)
// to plain C for ctypes.
#include <algorithm>
#include <cstdint>
#include <vector>

#include "eos.hpp"
#include "planet_builder.hpp"

namespace {
astraea::MaterialId mat(int m) { return static_cast<astraea::MaterialId>(m); }
}  // namespace

extern "C" {

void ref_eos_eval(int m, float rho, float u, float damage, float* p, float* cs, float* t, uint8_t* phase) {
    const astraea::eos::EosEvalResult r = astraea::eos::evaluateEosPsu(mat(m), rho, u, damage);
    *p = r.pressurePsu;
    *cs = r.soundSpeedPsu;
    *t = r.tempK;
    *phase = r.phaseId;
}

double ref_cold_energy_psu(int m, double rho) { return astraea::eos::computeColdEnergyPsu(mat(m), rho); }

void ref_eos_from_rho_temp_si(int m, double rho, double temp, double* u, double* p) {
    astraea::eos::evaluateFromRhoTempSI(mat(m), rho, temp, *u, *p);
}

double ref_invert_density_si(int m, double p, double temp, double guess) {
    return astraea::eos::invertDensityFromPressureTempSI(mat(m), p, temp, guess);
}

// PlanetBuilder::solveHydrostatic1D for an archetype (types.hpp:27-34) and mass; up to `cap`
// samples of r, rho, P, enclosed mass and T (any pointer may be NULL).  Returns the sample count.
int ref_solve_hydrostatic(int archetype, double mass_kg, double surface_temp_k, double core_temp_k, double* radius_m,
                          double* central_pressure_pa, int cap, double* r, double* rho, double* p, double* m,
                          double* t) {
    astraea::BodyInitSpec spec;
    spec.archetype = static_cast<astraea::PlanetArchetype>(archetype);
    spec.massKg = mass_kg;
    spec.surfaceTempK = surface_temp_k;
    spec.coreTempK = core_temp_k;
    const std::vector<astraea::RadialProfileSample> prof =
        astraea::PlanetBuilder::solveHydrostatic1D(spec, *radius_m, *central_pressure_pa);
    const int n = std::min<int>(cap, (int)prof.size());
    for (int k = 0; k < n; ++k) {
        if (r) r[k] = prof[k].rM;
        if (rho) rho[k] = prof[k].rhoKgM3;
        if (p) p[k] = prof[k].pressurePa;
        if (m) m[k] = prof[k].enclosedMassKg;
        if (t) t[k] = prof[k].tempK;
    }
    return (int)prof.size();
}

}  // extern "C"
