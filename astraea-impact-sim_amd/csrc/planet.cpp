// planet.cpp — layered planets in hydrostatic equilibrium with the simulation's own EOS.
//
// The reference's PlanetBuilder (engine/planet_builder.cpp) integrates dP/dr = -G m rho / r^2,
// dm/dr = 4 pi r^2 rho outward from a central pressure with RK4, rho from inverting the EOS at
// the pressure and at a geotherm temperature linear in pressure (20-46), switches material where
// the enclosed mass reaches each layer's cumulative mass fraction, stops at the surface pressure
// (56-264), and shoots on log P_c until the mass is reached (398-471).  Particles are then set on
// Fibonacci-sphere shells about one spacing (m / rho)^(1/3) thick, counts per shell in proportion
// to the shell's mass (largest remainder), each shell turned by a random rotation, with
// h0 = 1.2 (m / rho)^(1/3) and u from (rho, T) (474-735).
//
// Written here around the EOS of eos.hpp, so a particle's density and energy give back the
// profile's pressure when the simulation evaluates them.  The random rotations and failure
// strains come from this library's splitmix64 stream, not from std::mt19937_64.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <limits>
#include <numeric>
#include <vector>

#include "../../include/astraea_hip.h"
#include "eos.hpp"
#include "errors.hpp"

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kGsi = 6.67430e-11;  // engine/units.hpp:5
constexpr double kLUnit = 1.0e6, kMUnit = 1.0e24;
constexpr int kHHe = 3;

struct Sample {
    double r, rho, p, m, t;
    int mat;
};

uint64_t nextBits(uint64_t& s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
double uniformOpen(uint64_t& s) {  // (0, 1)
    return ((double)(nextBits(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

class Planet {
public:
    explicit Planet(const astraea_planet_spec& s) : spec(s) {
        const double fsum = std::accumulate(s.layer_mass_frac, s.layer_mass_frac + s.num_layers, 0.0);
        double acc = 0.0;
        for (uint32_t k = 0; k < s.num_layers; ++k) {
            acc += s.layer_mass_frac[k] / fsum;
            layerEnd[k] = (k + 1 == s.num_layers) ? std::numeric_limits<double>::infinity() : acc * s.mass_kg;
        }
        gasGiant = s.layer_material[s.num_layers - 1] == kHHe;
    }

    // Pressure (Pa) and u (J/kg) of material `mat` at density rho (kg/m^3) and temperature t (K):
    // u = u_cold + c_v T, P from the EOS at that u (eos.cu:390-442).
    void stateAt(int mat, double rho, double t, double& u, double& p) const {
        const astraea_hip::Material m = astraea_hip::material(mat);
        const double rhoPsu = rho / astraea_hip::kRhoUnit;
        const double uPsu = astraea_hip::coldEnergyPsu(mat, rhoPsu) + m.cv / astraea_hip::kUUnit * std::max(t, 0.0);
        u = uPsu * astraea_hip::kUUnit;
        p = astraea_hip::evaluateEos(mat, rhoPsu, uPsu, 0.0, spec.gamma_hhe).p * astraea_hip::kPUnit;
    }

    // Density with pressure p at temperature t (eos.cu:444-520): closed form for H/He, else a
    // bracket grown from the guess and false position (Illinois) in log rho.
    double densityAt(int mat, double p, double t, double guess) const {
        const astraea_hip::Material m = astraea_hip::material(mat);
        if (mat == kHHe) {
            if (p <= 0.0) return 1.0e-3;
            const double b = (spec.gamma_hhe - 1.0) * m.cv * std::max(t, 0.0);
            return std::max(2.0 * p / std::max(b + std::sqrt(b * b + 4.0 * m.kH2 * p), 1.0e-30), 1.0e-3);
        }
        auto f = [&](double rho) {
            double u, q;
            stateAt(mat, rho, t, u, q);
            return q - p;
        };
        double lo = std::clamp(guess > 0.0 ? guess : m.rho0, 0.05 * m.rho0, 50.0 * m.rho0), hi = lo;
        if (f(lo) > 0.0) {
            for (int i = 0; i < 80 && f(lo) > 0.0 && lo > 1.0e-6; ++i) lo = std::max(lo * 0.7, 1.0e-6);
            if (f(lo) > 0.0) return lo;
        } else {
            for (int i = 0; i < 80 && f(hi) < 0.0 && hi < 1.0e8; ++i) hi *= 1.4;
            if (f(hi) < 0.0) return hi;
        }
        // Illinois false position in log rho on the bracket [lo, hi], f(lo) <= 0 < f(hi)
        double a = std::log(lo), b = std::log(hi), fa = f(lo), fb = f(hi);
        int side = 0;
        for (int it = 0; it < 200; ++it) {
            const double c = (fb - fa) != 0.0 ? b - fb * (b - a) / (fb - fa) : 0.5 * (a + b);
            const double fc = f(std::exp(c));
            if (std::abs(fc) <= 1.0e-12 * std::max(1.0, std::abs(p)) || b - a <= 1.0e-15) return std::exp(c);
            if (fc < 0.0) {
                a = c; fa = fc;
                if (side == -1) fb *= 0.5;
                side = -1;
            } else {
                b = c; fb = fc;
                if (side == 1) fa *= 0.5;
                side = 1;
            }
        }
        return std::exp(0.5 * (a + b));
    }

    double temperatureAt(double p) const {  // planet_builder.cpp:20-29
        const double frac = std::clamp((p - spec.surface_pressure_pa) / std::max(pc - spec.surface_pressure_pa, 1.0), 0.0, 1.0);
        return spec.surface_temp_k + (spec.core_temp_k - spec.surface_temp_k) * frac;
    }

    // One outward integration from central pressure `pcIn`; returns the mass at the surface
    // (or above 20 M when the mass runs away) and records the profile when asked.
    double integrate(double pcIn, double dr, bool record, bool& runaway) {
        pc = pcIn;
        runaway = false;
        profile.clear();
        const double ps = spec.surface_pressure_pa;
        uint32_t k = 0;
        int mat = spec.layer_material[0];
        double guess = astraea_hip::material(mat).rho0;
        auto rhoOf = [&](int mt, double p) {
            const double r = densityAt(mt, std::max(p, ps), temperatureAt(p), guess);
            guess = r;
            return r;
        };
        const double rhoC = rhoOf(mat, pc);
        auto deriv = [&](double r, double p, double m, double& dp, double& dm) {
            const double rho = rhoOf(mat, p);
            dm = 4.0 * kPi * r * r * rho;
            dp = r > 0.0 ? -kGsi * m * rho / (r * r) : 0.0;
        };
        auto rk4 = [&](double r, double p, double m, double h, double& p1, double& m1) {
            double a1, b1, a2, b2, a3, b3, a4, b4;
            deriv(r, p, m, a1, b1);
            deriv(r + 0.5 * h, std::max(ps, p + 0.5 * h * a1), m + 0.5 * h * b1, a2, b2);
            deriv(r + 0.5 * h, std::max(ps, p + 0.5 * h * a2), m + 0.5 * h * b2, a3, b3);
            deriv(r + h, std::max(ps, p + h * a3), m + h * b3, a4, b4);
            p1 = p + h / 6.0 * (a1 + 2.0 * a2 + 2.0 * a3 + a4);
            m1 = m + h / 6.0 * (b1 + 2.0 * b2 + 2.0 * b3 + b4);
        };
        // first step from the centre: uniform-density sphere
        double h0 = (gasGiant && k + 1 < spec.num_layers) ? 0.25 * dr : dr;
        h0 = std::min(h0, std::cbrt(3.0 * std::min(layerEnd[0], spec.mass_kg) / (4.0 * kPi * rhoC)));
        double r = h0;
        double p = pc - 2.0 / 3.0 * kPi * kGsi * rhoC * rhoC * h0 * h0;
        double m = 4.0 / 3.0 * kPi * rhoC * h0 * h0 * h0;
        if (record) {
            profile.push_back({0.0, rhoC, pc, 0.0, temperatureAt(pc), mat});
            if (p > ps) profile.push_back({r, rhoOf(mat, p), p, m, temperatureAt(p), mat});
        }
        for (int step = 0; step < 200000; ++step) {
            if (p <= ps) break;
            if (m > 20.0 * spec.mass_kg) {
                runaway = true;
                return m;
            }
            const double h = (gasGiant && k + 1 < spec.num_layers) ? 0.25 * dr : dr;
            double p1, m1;
            rk4(r, p, m, h, p1, m1);
            const bool surface = p1 <= ps;
            const bool layer = k + 1 < spec.num_layers && m1 >= layerEnd[k];
            if (!surface && !layer) {
                r += h; p = p1; m = m1;
                if (record) profile.push_back({r, rhoOf(mat, p), p, m, temperatureAt(p), mat});
                continue;
            }
            // the sub-step to the first event: surface pressure or the layer's end mass
            double lo = 0.0, hi = h;
            for (int it = 0; it < 60; ++it) {
                const double mid = 0.5 * (lo + hi);
                rk4(r, p, m, mid, p1, m1);
                const bool past = p1 <= ps || (k + 1 < spec.num_layers && m1 >= layerEnd[k]);
                if (past) hi = mid; else lo = mid;
            }
            rk4(r, p, m, hi, p1, m1);
            r += hi;
            if (p1 <= ps) {  // surface
                p = ps; m = m1;
                if (record) profile.push_back({r, densityAt(mat, ps, spec.surface_temp_k, guess), ps, m, spec.surface_temp_k, mat});
                break;
            }
            p = p1; m = layerEnd[k];
            if (record) profile.push_back({r, rhoOf(mat, p), p, m, temperatureAt(p), mat});
            ++k;
            mat = spec.layer_material[k];
            guess = astraea_hip::material(mat).rho0;
            if (record) profile.push_back({r, rhoOf(mat, p), p, m, temperatureAt(p), mat});
        }
        radius = r;
        return m;
    }

    // Central pressure for the spec's mass (planet_builder.cpp:398-471).
    bool solve() {
        double invRho = 0.0;
        for (uint32_t k = 0; k < spec.num_layers; ++k) {
            const int mat = spec.layer_material[k];
            invRho += spec.layer_mass_frac[k] / (mat == kHHe ? 600.0 : astraea_hip::material(mat).rho0);
        }
        double rEst = std::cbrt(3.0 * spec.mass_kg * invRho / (4.0 * kPi));
        if (gasGiant) rEst = std::max(rEst, 6.5e7);
        dr = std::max(rEst / 800.0, 100.0);
        const double pEst = 3.0 * kGsi * spec.mass_kg * spec.mass_kg / (8.0 * kPi * std::pow(rEst, 4.0));
        double lo = std::log(std::max(spec.surface_pressure_pa + 1.0e4, 1.0e-3 * pEst));
        double hi = std::log(std::max(spec.surface_pressure_pa + 1.0e6, 1.0e3 * pEst));
        bool run = false;
        for (int i = 0; i < 20 && (integrate(std::exp(lo), dr, false, run) > spec.mass_kg || run); ++i) lo -= 2.0;
        for (int i = 0; i < 20 && !(integrate(std::exp(hi), dr, false, run) >= spec.mass_kg || run); ++i) hi += 2.0;
        double mid = 0.5 * (lo + hi);
        for (int it = 0; it < 200 && hi - lo > 1.0e-13; ++it) {
            mid = 0.5 * (lo + hi);
            const double m = integrate(std::exp(mid), dr, false, run);
            if (!run && std::abs(m - spec.mass_kg) <= 1.0e-7 * spec.mass_kg) break;
            if (run || m > spec.mass_kg) hi = mid; else lo = mid;
        }
        const double m = integrate(std::exp(mid), dr, true, run);
        mass = m;
        return !run && profile.size() >= 2 && std::isfinite(m) && std::abs(m - spec.mass_kg) <= 1.0e-3 * spec.mass_kg;
    }

    // Profile sample at enclosed mass mq (linear in mass between samples).
    Sample atMass(double mq) const {
        auto it = std::lower_bound(profile.begin(), profile.end(), mq, [](const Sample& s, double v) { return s.m < v; });
        if (it == profile.begin()) return profile.front();
        if (it == profile.end()) return profile.back();
        const Sample& b = *it;
        const Sample& a = *(it - 1);
        const double t = b.m > a.m ? (mq - a.m) / (b.m - a.m) : 0.0;
        return {a.r + t * (b.r - a.r), a.rho + t * (b.rho - a.rho), a.p + t * (b.p - a.p), mq,
                a.t + t * (b.t - a.t), a.mat};
    }
    double massAtRadius(double rq) const {
        auto it = std::lower_bound(profile.begin(), profile.end(), rq, [](const Sample& s, double v) { return s.r < v; });
        if (it == profile.begin()) return profile.front().m;
        if (it == profile.end()) return profile.back().m;
        const Sample& b = *it;
        const Sample& a = *(it - 1);
        const double t = b.r > a.r ? (rq - a.r) / (b.r - a.r) : 0.0;
        return a.m + t * (b.m - a.m);
    }

    const astraea_planet_spec spec;
    std::array<double, ASTRAEA_MAX_LAYERS> layerEnd{};
    bool gasGiant = false;
    double pc = 0.0, dr = 0.0, radius = 0.0, mass = 0.0;
    std::vector<Sample> profile;
};

// Uniformly random rotation from a unit quaternion (Shoemake's method).
std::array<double, 9> randomRotation(uint64_t& s) {
    const double u1 = uniformOpen(s), u2 = uniformOpen(s), u3 = uniformOpen(s);
    const double a = std::sqrt(1.0 - u1), b = std::sqrt(u1);
    const double w = a * std::sin(2.0 * kPi * u2), x = a * std::cos(2.0 * kPi * u2);
    const double y = b * std::sin(2.0 * kPi * u3), z = b * std::cos(2.0 * kPi * u3);
    return {1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
            2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
            2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)};
}

bool validSpec(const astraea_planet_spec* s) {
    if (!s || !(s->mass_kg > 0.0) || s->num_layers < 1 || s->num_layers > ASTRAEA_MAX_LAYERS) return false;
    double fsum = 0.0;
    for (uint32_t k = 0; k < s->num_layers; ++k) {
        if (s->layer_material[k] > 4 || !(s->layer_mass_frac[k] >= 0.0)) return false;
        if (s->layer_material[k] == kHHe && k + 1 != s->num_layers) return false;
        fsum += s->layer_mass_frac[k];
    }
    return fsum > 0.0 && s->surface_temp_k >= 0.0 && s->core_temp_k >= 0.0 && s->gamma_hhe > 1.0 &&
           s->h_factor > 0.0 && s->surface_pressure_pa >= 0.0;
}

}  // namespace

extern "C" {

void astraea_default_planet_spec(astraea_planet_spec* s) {
    if (!s) return;
    *s = astraea_planet_spec{};
    s->mass_kg = 5.97219e24;
    s->num_layers = 2;
    s->layer_material[0] = 0;
    s->layer_material[1] = 1;
    s->layer_mass_frac[0] = 0.32;
    s->layer_mass_frac[1] = 0.68;
    s->surface_temp_k = 288.0;
    s->core_temp_k = 5500.0;
    s->surface_pressure_pa = 0.0;
    s->gamma_hhe = 1.4;
    s->h_factor = 1.2;
    s->seed = 42;
}

int astraea_planet_profile(const astraea_planet_spec* spec, uint32_t cap, double* r, double* rho, double* p, double* m,
                           double* t, uint8_t* mat, astraea_planet_info* info) {
    if (!validSpec(spec)) return astraea_hip::reportError(ASTRAEA_E_INVALID, "invalid planet spec");
    Planet pl(*spec);
    if (!pl.solve())
        return astraea_hip::reportError(ASTRAEA_E_INVALID, "no hydrostatic profile reaches %g kg", spec->mass_kg);
    const uint32_t n = (uint32_t)std::min<size_t>(cap, pl.profile.size());
    for (uint32_t i = 0; i < n; ++i) {
        const Sample& s = pl.profile[i];
        if (r) r[i] = s.r;
        if (rho) rho[i] = s.rho;
        if (p) p[i] = s.p;
        if (m) m[i] = s.m;
        if (t) t[i] = s.t;
        if (mat) mat[i] = (uint8_t)s.mat;
    }
    if (info) *info = {pl.mass, pl.radius, pl.pc, (uint32_t)pl.profile.size()};
    return ASTRAEA_OK;
}

int astraea_build_planet(const astraea_planet_spec* spec, uint32_t n, astraea_body_out* out, float* tempK,
                         astraea_planet_info* info) {
    if (!validSpec(spec) || !out || n == 0)
        return astraea_hip::reportError(ASTRAEA_E_INVALID, "invalid planet spec, output or particle count");
    Planet pl(*spec);
    if (!pl.solve())
        return astraea_hip::reportError(ASTRAEA_E_INVALID, "no hydrostatic profile reaches %g kg", spec->mass_kg);
    const uint32_t L = spec->num_layers;
    const double M = pl.mass;
    // particles per layer in proportion to the mass fractions, at least one each (474-509)
    const double fsum = std::accumulate(spec->layer_mass_frac, spec->layer_mass_frac + L, 0.0);
    std::vector<uint32_t> nk(L, 0);
    uint32_t given = 0;
    for (uint32_t k = 0; k + 1 < L; ++k) {
        uint32_t c = (uint32_t)std::llround(spec->layer_mass_frac[k] / fsum * n);
        if (n >= L && c == 0) c = 1;
        const uint32_t room = n - given - (n >= L ? L - 1 - k : 0u);
        nk[k] = std::min(c, room);
        given += nk[k];
    }
    nk[L - 1] = n - given;
    std::vector<double> mEnd(L);
    double acc = 0.0;
    for (uint32_t k = 0; k < L; ++k) {
        acc += spec->layer_mass_frac[k] / fsum;
        mEnd[k] = k + 1 == L ? M : acc * M;
    }
    const double mNominal = M / n;
    const double golden = kPi * (3.0 - std::sqrt(5.0));
    uint64_t rng = spec->seed ^ 0x2545F4914F6CDD1DULL;
    uint32_t w = 0;
    double mStart = 0.0;
    for (uint32_t k = 0; k < L; ++k) {
        const double layerMass = std::max(mEnd[k] - mStart, 0.0);
        if (nk[k] == 0) {
            mStart = mEnd[k];
            continue;
        }
        const int mat = spec->layer_material[k];
        const double mP = layerMass / nk[k];
        const Sample sIn = pl.atMass(mStart), sOut = pl.atMass(mEnd[k]);
        // shell boundaries in enclosed mass, about one particle spacing apart (560-627)
        std::vector<double> bounds{mStart};
        double curM = mStart, curR = sIn.r;
        if (k == 0 && nk[k] >= 2) {  // one particle at the centre
            const double cM = std::min(mStart + mP, mEnd[k]);
            const double cR = std::cbrt(3.0 * cM / (4.0 * kPi * std::max(sIn.rho, 1.0e-3)));
            if (cM < mEnd[k] && cR < sOut.r) {
                bounds.push_back(cM);
                curM = cM;
                curR = cR;
            }
        }
        while (curM < mEnd[k]) {
            const double spacing = std::cbrt(mNominal / std::max(pl.atMass(curM).rho, 1.0e-3));
            const double step = (k == 0 && nk[k] >= 64 && curR < 3.5 * spacing) ? 0.62 * spacing : spacing;
            const double nextR = curR + step;
            if (nextR >= sOut.r - 0.35 * step) break;
            const double nextM = std::clamp(pl.massAtRadius(nextR), curM + 1.0e-9 * layerMass, mEnd[k]);
            if (nextM >= mEnd[k]) break;
            bounds.push_back(nextM);
            curM = nextM;
            curR = nextR;
        }
        bounds.push_back(mEnd[k]);
        if (bounds.size() == 2 && nk[k] >= 128 && sOut.r > sIn.r) {  // a thin crust: three sub-shells
            const double sp = std::cbrt(mNominal / std::max(0.5 * (sIn.rho + sOut.rho), 1.0e-3));
            if (sOut.r - sIn.r > 0.25 * sp) {
                bounds = {mStart};
                for (int sub = 1; sub < 3; ++sub) {
                    const double mq = pl.massAtRadius(sIn.r + sub / 3.0 * (sOut.r - sIn.r));
                    if (mq > bounds.back() && mq < mEnd[k]) bounds.push_back(mq);
                }
                bounds.push_back(mEnd[k]);
            }
        }
        // particles per shell: largest remainder (629-650)
        const size_t S = bounds.size() - 1;
        std::vector<uint32_t> cnt(S);
        std::vector<double> rem(S);
        uint32_t sum = 0;
        for (size_t s = 0; s < S; ++s) {
            const double exact = layerMass > 0.0 ? nk[k] * (bounds[s + 1] - bounds[s]) / layerMass : 0.0;
            cnt[s] = (uint32_t)std::floor(exact);
            rem[s] = exact - cnt[s];
            sum += cnt[s];
        }
        if (sum < nk[k]) {
            std::vector<size_t> idx(S);
            std::iota(idx.begin(), idx.end(), 0);
            std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return rem[a] > rem[b]; });
            for (uint32_t q = 0; q < nk[k] - sum; ++q) cnt[idx[q % S]] += 1;
        }
        const astraea_hip::Material mp = astraea_hip::material(mat);
        for (size_t s = 0; s < S; ++s) {
            if (cnt[s] == 0) continue;
            const Sample sm = pl.atMass(0.5 * (bounds[s] + bounds[s + 1]));
            const double rPsu = (s == 0 && k == 0 && cnt[s] == 1) ? 0.0 : sm.r / kLUnit;
            const double rho = std::max(sm.rho, 1.0e-3);
            double u, p;
            pl.stateAt(mat, rho, sm.t, u, p);
            const double rhoPsu = rho / astraea_hip::kRhoUnit;
            const double h0 = spec->h_factor * std::cbrt(mP / kMUnit / rhoPsu);
            const std::array<double, 9> R = randomRotation(rng);
            for (uint32_t q = 0; q < cnt[s]; ++q, ++w) {
                double x = 0.0, y = 0.0, z = 0.0;
                if (rPsu > 0.0) {
                    const double cz = cnt[s] == 1 ? 1.0 : 1.0 - 2.0 * (q + 0.5) / cnt[s];
                    const double rxy = std::sqrt(std::max(0.0, 1.0 - cz * cz));
                    const double a = golden * q;
                    const double ux = rxy * std::cos(a), uy = rxy * std::sin(a);
                    x = rPsu * (R[0] * ux + R[1] * uy + R[2] * cz);
                    y = rPsu * (R[3] * ux + R[4] * uy + R[5] * cz);
                    z = rPsu * (R[6] * ux + R[7] * uy + R[8] * cz);
                }
                // failure strain: eps_min clamp((-ln U)^(1/9), 1, 3) (planet_builder.cpp:691-694)
                const double wd = std::max(-std::log(uniformOpen(rng)), 1.0e-6);
                const double eps = mp.epsMin * std::clamp(std::pow(wd, 1.0 / 9.0), 1.0, 3.0);
                if (out->x) out->x[w] = (float)x;
                if (out->y) out->y[w] = (float)y;
                if (out->z) out->z[w] = (float)z;
                if (out->m) out->m[w] = (float)(mP / kMUnit);
                if (out->h) out->h[w] = (float)h0;
                if (out->rho) out->rho[w] = (float)rhoPsu;
                if (out->u) out->u[w] = (float)(u / astraea_hip::kUUnit);
                if (out->eps_fail) out->eps_fail[w] = (float)eps;
                if (out->mat) out->mat[w] = (uint8_t)mat;
                if (tempK) tempK[w] = (float)sm.t;
            }
        }
        mStart = mEnd[k];
    }
    if (w != n) return astraea_hip::reportError(ASTRAEA_E_INVALID, "placed %u of %u particles", w, n);
    if (info) *info = {M, pl.radius, pl.pc, (uint32_t)pl.profile.size()};
    return ASTRAEA_OK;
}

void astraea_cube_face_uv_to_normal(uint32_t face, float u, float v, float* nx, float* ny, float* nz) {
    const float sx = 2.0f * u - 1.0f, sy = 2.0f * v - 1.0f;
    // face axes (planet_builder.cpp:329-349): +X, -X, +Y, -Y, +Z, -Z
    float vx, vy, vz;
    switch (face % 6u) {
        case 0: vx = 1.0f; vy = sy; vz = -sx; break;
        case 1: vx = -1.0f; vy = sy; vz = sx; break;
        case 2: vx = sx; vy = 1.0f; vz = -sy; break;
        case 3: vx = sx; vy = -1.0f; vz = sy; break;
        case 4: vx = sx; vy = sy; vz = 1.0f; break;
        default: vx = -sx; vy = sy; vz = -1.0f; break;
    }
    const float inv = 1.0f / std::sqrt(vx * vx + vy * vy + vz * vz);
    *nx = vx * inv;
    *ny = vy * inv;
    *nz = vz * inv;
}

void astraea_normal_to_cube_face_uv(float nx, float ny, float nz, uint32_t* face, float* u, float* v) {
    // the dominant axis picks the face; the other two, divided by it, are the face coordinates
    // (planet_builder.cpp:351-395)
    const float ax = std::fabs(nx), ay = std::fabs(ny), az = std::fabs(nz);
    float sx, sy;
    if (ax >= ay && ax >= az) {
        const float d = std::max(ax, 1.0e-12f);
        *face = nx >= 0.0f ? 0u : 1u;
        sx = (nx >= 0.0f ? -nz : nz) / d;
        sy = ny / d;
    } else if (ay >= ax && ay >= az) {
        const float d = std::max(ay, 1.0e-12f);
        *face = ny >= 0.0f ? 2u : 3u;
        sx = nx / d;
        sy = (ny >= 0.0f ? -nz : nz) / d;
    } else {
        const float d = std::max(az, 1.0e-12f);
        *face = nz >= 0.0f ? 4u : 5u;
        sx = (nz >= 0.0f ? nx : -nx) / d;
        sy = ny / d;
    }
    *u = std::clamp(0.5f * (sx + 1.0f), 0.0f, 1.0f);
    *v = std::clamp(0.5f * (sy + 1.0f), 0.0f, 1.0f);
}

int astraea_select_tracers(uint32_t n, const float* x, const float* y, const float* z, const float* h0,
                           const uint8_t* mat, uint32_t gridRes, int32_t* tracerIndex, float* restRadius,
                           float* restNormal, uint8_t* ocean, uint32_t* outCount) {
    if (!x || !y || !z || !h0 || !mat || !tracerIndex || !restRadius || !restNormal || !ocean || !outCount)
        return astraea_hip::reportError(ASTRAEA_E_INVALID, "null argument");
    constexpr uint8_t kWaterIce = 2;
    for (uint32_t i = 0; i < n; ++i) {
        tracerIndex[i] = -1;
        restRadius[i] = 0.0f;
        restNormal[3 * i] = restNormal[3 * i + 1] = restNormal[3 * i + 2] = 0.0f;
        ocean[i] = 0;
    }
    *outCount = 0;
    const uint64_t target64 = 6ull * gridRes * gridRes;
    if (gridRes == 0 || target64 > 65535ull || n < target64) return ASTRAEA_OK;
    const uint32_t target = (uint32_t)target64;
    // candidates from the outside in (planet_builder.cpp:748-771)
    std::vector<float> rad(n);
    std::vector<uint32_t> cand;
    for (uint32_t i = 0; i < n; ++i) {
        rad[i] = std::sqrt(x[i] * x[i] + y[i] * y[i] + z[i] * z[i]);
        if (rad[i] > 1.0e-6f) cand.push_back(i);
    }
    std::stable_sort(cand.begin(), cand.end(), [&](uint32_t a, uint32_t b) { return rad[a] > rad[b]; });
    for (uint32_t i = 0; i < n && cand.size() < target; ++i)
        if (rad[i] <= 1.0e-6f) cand.push_back(i);
    if (cand.empty()) return ASTRAEA_OK;
    // the pool: the outermost shell (1.8 smoothing lengths) and any ocean, at least 3 x target
    const float rMax = rad[cand.front()], drSurf = h0[cand.front()];
    const size_t minPool = std::min(cand.size(), (size_t)target * 3u);
    size_t pool = 0;
    for (; pool < cand.size(); ++pool) {
        const uint32_t q = cand[pool];
        const bool shell = rad[q] >= rMax - 1.8f * drSurf || mat[q] == kWaterIce;
        if (!shell && pool >= minPool) break;
    }
    if (pool < target) pool = std::min(cand.size(), (size_t)target);
    cand.resize(pool);
    std::vector<float> cn(3 * pool);
    std::vector<uint8_t> used(pool, 0);
    std::vector<std::vector<uint32_t>> bins(target);
    const uint32_t G = gridRes;
    for (size_t k = 0; k < pool; ++k) {
        const uint32_t q = cand[k];
        if (rad[q] > 1.0e-6f) {
            const float inv = 1.0f / rad[q];
            cn[3 * k] = x[q] * inv; cn[3 * k + 1] = y[q] * inv; cn[3 * k + 2] = z[q] * inv;
        } else {
            cn[3 * k] = 0.0f; cn[3 * k + 1] = 0.0f; cn[3 * k + 2] = 1.0f;
        }
        uint32_t face;
        float u, v;
        astraea_normal_to_cube_face_uv(cn[3 * k], cn[3 * k + 1], cn[3 * k + 2], &face, &u, &v);
        const uint32_t gx = std::min(G - 1u, (uint32_t)(u * (float)G));
        const uint32_t gy = std::min(G - 1u, (uint32_t)(v * (float)G));
        bins[face * G * G + gy * G + gx].push_back((uint32_t)k);
    }
    std::vector<float> tn(3 * (size_t)target);
    std::vector<uint8_t> filled(target, 0);
    uint32_t assigned = 0;
    auto assign = [&](uint32_t b, size_t k) {
        const uint32_t q = cand[k];
        used[k] = 1;
        filled[b] = 1;
        ++assigned;
        tracerIndex[q] = (int32_t)b;
        restRadius[q] = rad[q];
        for (int a = 0; a < 3; ++a) restNormal[3 * q + a] = cn[3 * k + a];
        ocean[q] = mat[q] == kWaterIce ? 1u : 0u;
    };
    auto best = [&](uint32_t b, const std::vector<uint32_t>* within) {
        float bestDot = -2.0f;
        size_t bestK = SIZE_MAX;
        auto consider = [&](size_t k) {
            if (used[k]) return;
            const float d = cn[3 * k] * tn[3 * b] + cn[3 * k + 1] * tn[3 * b + 1] + cn[3 * k + 2] * tn[3 * b + 2];
            if (d > bestDot) {
                bestDot = d;
                bestK = k;
            }
        };
        if (within)
            for (uint32_t k : *within) consider(k);
        else
            for (size_t k = 0; k < pool; ++k) consider(k);
        return bestK;
    };
    // pass A: in each cube-sphere cell, the candidate nearest the cell centre (planet_builder.cpp:834-859)
    for (uint32_t face = 0; face < 6; ++face)
        for (uint32_t gy = 0; gy < G; ++gy)
            for (uint32_t gx = 0; gx < G; ++gx) {
                const uint32_t b = face * G * G + gy * G + gx;
                astraea_cube_face_uv_to_normal(face, (gx + 0.5f) / (float)G, (gy + 0.5f) / (float)G, &tn[3 * b],
                                               &tn[3 * b + 1], &tn[3 * b + 2]);
                const size_t k = best(b, &bins[b]);
                if (k != SIZE_MAX) assign(b, k);
            }
    // pass B: empty cells take the nearest unused candidate of the pool (861-881)
    for (uint32_t b = 0; b < target; ++b) {
        if (filled[b]) continue;
        const size_t k = best(b, nullptr);
        if (k != SIZE_MAX) assign(b, k);
    }
    if (assigned < target) {  // compact the indices (883-891)
        int32_t next = 0;
        for (uint32_t i = 0; i < n; ++i)
            if (tracerIndex[i] >= 0) tracerIndex[i] = next++;
    }
    *outCount = assigned;
    return ASTRAEA_OK;
}

}  // extern "C"
