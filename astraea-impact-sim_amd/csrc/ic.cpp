// ic.cpp — host-side initial conditions for the benchmark bodies.
//
// The reference builds bodies from a 1D hydrostatic RK4 profile and Fibonacci
// shells (engine/planet_builder.cpp:398-906).  The benchmark bodies instead
// sit on a seed-jittered hexagonal-close-packed lattice with uniform density
// per layer, so the hydrostatic pressure of piecewise-constant density has a
// closed form; u is then the root of P_EOS(rho_layer, u) = P(r).  Two bodies
// are put on the Keplerian approach orbit of Simulation::loadScenario
// (engine/simulation.cpp:1389-1436) and shifted to zero COM and momentum
// (simulation.cpp:919-944).
//
// A body whose outermost layer is H/He gets a gas envelope instead: the cold part of the H/He
// EOS is the n = 1 polytrope P = K rho^2 (eos.cu:331), so the envelope in hydrostatic equilibrium
// around uniform-density core layers solves d rho / dr = -G M(r) / (2 K r^2).  The total mass and
// the density at the core's top are found by shooting, so that the envelope holds its mass
// fraction and ends (rho = 0) at the requested radius.  Particles have equal masses; a uniform
// lattice ball is stretched radially so that each particle's rank in radius matches its
// enclosed mass, and u is the cold energy K rho (no thermal part, P = K rho^2 exactly).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/astraea_hip.h"
#include "eos.hpp"

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kGsi = 6.67430e-11;   // engine/units.hpp:5
constexpr double kLUnit = 1.0e6, kMUnit = 1.0e24;

uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
// uniform in [0, 1) with 53 random bits
double uniform01(uint64_t& s) { return (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

struct Layer {
    double rIn, rOut, rho, mEncIn;
    int mat;
};

// Hydrostatic pressure of piecewise-constant density, P(R) = 0:
// within a layer, M(r) = M_in + (4/3) pi rho (r^3 - a^3) and
// P(r) = P(b) + G rho [ (M_in - (4/3) pi rho a^3)(1/r - 1/b) + (2/3) pi rho (b^2 - r^2) ].
double pressureAt(const std::vector<Layer>& L, const std::vector<double>& pTop, double r) {
    for (size_t k = 0; k < L.size(); ++k) {
        if (r <= L[k].rOut || k + 1 == L.size()) {
            const Layer& l = L[k];
            const double kk = 4.0 / 3.0 * kPi * l.rho;
            const double c = l.mEncIn - kk * l.rIn * l.rIn * l.rIn;
            const double rr = std::max(r, 1.0e-9 * l.rOut);
            const double inner = (c != 0.0) ? c * (1.0 / rr - 1.0 / l.rOut) : 0.0;
            const double p = pTop[k] + kGsi * l.rho * (inner + 0.5 * kk * (l.rOut * l.rOut - rr * rr));
            return std::max(p, 0.0);
        }
    }
    return 0.0;
}

// u (PSU) with P_EOS(rho, u) = p (Pa), bisection; u >= u_cold + cv * T_surface.
double energyForPressure(int mat, double rhoPsu, double pSi, double tSurface, double gammaHHe) {
    const astraea_hip::Material m = astraea_hip::material(mat);
    const double uc = astraea_hip::coldEnergyPsu(mat, rhoPsu);
    const double cv = m.cv / astraea_hip::kUUnit;
    double lo = uc + cv * tSurface;
    const double target = pSi / astraea_hip::kPUnit;
    auto P = [&](double u) { return astraea_hip::evaluateEos(mat, rhoPsu, u, 0.0, gammaHHe).p; };
    if (P(lo) >= target) return lo;
    double hi = lo + 1.0;
    int guard = 0;
    while (P(hi) < target && guard++ < 200) hi = lo + 2.0 * (hi - lo);
    for (int it = 0; it < 200; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (P(mid) < target) lo = mid; else hi = mid;
        if (hi - lo <= 1.0e-15 * hi) break;
    }
    return 0.5 * (lo + hi);
}

struct LatticePoint {
    double x, y, z, r2;
    uint64_t id;
};

// The n points of a seed-jittered HCP lattice of spacing s nearest to the origin, re-centred on
// their centroid, in lattice order.  Returns false if the over-generated ball holds fewer than n.
bool latticeBall(uint32_t n, double R, double jitter, uint64_t seed, std::vector<LatticePoint>& pts, double& s) {
    const double V = 4.0 / 3.0 * kPi * R * R * R;
    s = std::cbrt(std::sqrt(2.0) * V / (double)n);
    const double dzL = s * std::sqrt(2.0 / 3.0), dyR = s * std::sqrt(3.0) / 2.0;
    const double Rg = 1.15 * R + 2.0 * s;
    const int kz = (int)std::ceil(Rg / dzL) + 1, jy = (int)std::ceil(Rg / dyR) + 1, ix = (int)std::ceil(Rg / s) + 1;
    uint64_t rng = seed * 0x2545F4914F6CDD1DULL + 0x9E3779B97F4A7C15ULL;
    pts.clear();
    pts.reserve((size_t)(1.6 * n) + 64);
    const double jit = jitter * s;
    uint64_t id = 0;
    for (int k = -kz; k <= kz; ++k) {
        const bool odd = (k & 1) != 0;
        const double ox = odd ? 0.5 * s : 0.0, oy = odd ? s / (2.0 * std::sqrt(3.0)) : 0.0;
        const double z0 = k * dzL;
        for (int j = -jy; j <= jy; ++j) {
            const double y0 = j * dyR + oy;
            const double xo = ((j & 1) ? 0.5 * s : 0.0) + ox;
            for (int i = -ix; i <= ix; ++i) {
                const double x0 = i * s + xo;
                ++id;
                if (x0 * x0 + y0 * y0 + z0 * z0 > Rg * Rg) continue;
                uint64_t h = rng ^ (id * 0xD1B54A32D192ED03ULL);
                const double jx = (2.0 * uniform01(h) - 1.0) * jit;
                const double jy2 = (2.0 * uniform01(h) - 1.0) * jit;
                const double jz = (2.0 * uniform01(h) - 1.0) * jit;
                const double x = x0 + jx, y = y0 + jy2, z = z0 + jz;
                pts.push_back({x, y, z, x * x + y * y + z * z, id});
            }
        }
    }
    if (pts.size() < n) return false;
    std::nth_element(pts.begin(), pts.begin() + (n - 1), pts.end(), [](const LatticePoint& a, const LatticePoint& b) {
        return a.r2 < b.r2 || (a.r2 == b.r2 && a.id < b.id);
    });
    pts.resize(n);
    std::sort(pts.begin(), pts.end(), [](const LatticePoint& a, const LatticePoint& b) { return a.id < b.id; });
    double gx = 0.0, gy = 0.0, gz = 0.0;
    for (const LatticePoint& q : pts) { gx += q.x; gy += q.y; gz += q.z; }
    gx /= n; gy /= n; gz /= n;
    for (LatticePoint& q : pts) { q.x -= gx; q.y -= gy; q.z -= gz; q.r2 = q.x * q.x + q.y * q.y + q.z * q.z; }
    return true;
}

// Envelope of a gas giant: RK4 on (rho, M) outward from the core's top, SI units.
struct Envelope {
    std::vector<double> r, m, rho;  // samples from the core's top to the surface
    double rSurface = 0.0, mEnvelope = 0.0;
};

Envelope integrateEnvelope(double K, double rCore, double mCore, double rhoTop, double dr, bool keep) {
    Envelope e;
    double r = rCore, m = mCore, rho = rhoTop;
    auto f = [&](double rr, double rh, double mm, double& dRho, double& dM) {
        dRho = -kGsi * mm / (2.0 * K * rr * rr);
        dM = 4.0 * kPi * rr * rr * std::max(rh, 0.0);
    };
    if (keep) { e.r.push_back(r); e.m.push_back(m); e.rho.push_back(rho); }
    for (int it = 0; it < 4000000; ++it) {
        double a1, b1, a2, b2, a3, b3, a4, b4;
        f(r, rho, m, a1, b1);
        f(r + 0.5 * dr, rho + 0.5 * dr * a1, m + 0.5 * dr * b1, a2, b2);
        f(r + 0.5 * dr, rho + 0.5 * dr * a2, m + 0.5 * dr * b2, a3, b3);
        f(r + dr, rho + dr * a3, m + dr * b3, a4, b4);
        const double rn = rho + dr / 6.0 * (a1 + 2.0 * a2 + 2.0 * a3 + a4);
        const double mn = m + dr / 6.0 * (b1 + 2.0 * b2 + 2.0 * b3 + b4);
        if (rn <= 0.0) {
            const double t = rho / (rho - rn);
            e.rSurface = r + t * dr;
            e.mEnvelope = m + t * (mn - m) - mCore;
            if (keep) { e.r.push_back(e.rSurface); e.m.push_back(m + t * (mn - m)); e.rho.push_back(0.0); }
            return e;
        }
        rho = rn; m = mn; r += dr;
        if (keep) { e.r.push_back(r); e.m.push_back(m); e.rho.push_back(rho); }
    }
    e.rSurface = r;
    e.mEnvelope = m - mCore;
    return e;
}

// Core layers of uniform density under an n = 1 polytrope H/He envelope (P = K rho^2) holding the
// outermost layer's mass fraction: for a total mass M the core radii follow from the densities,
// the density at the core's top from the envelope's mass, and the surface from the envelope.
struct GasGiantModel {
    int nl = 0, nc = 0;
    double K = 0.0, fEnv = 0.0, dr = 0.0;
    double rho[ASTRAEA_MAX_LAYERS] = {}, f[ASTRAEA_MAX_LAYERS] = {};

    bool init(const astraea_body_spec* spec) {
        nl = (int)spec->num_layers;
        nc = nl - 1;
        K = astraea_hip::material(3).kH2;
        double fsum = 0.0;
        for (int k = 0; k < nl; ++k) fsum += spec->layer_mass_frac[k];
        if (!(fsum > 0.0)) return false;
        for (int k = 0; k < nl; ++k) {
            f[k] = spec->layer_mass_frac[k] / fsum;
            rho[k] = spec->layer_density_kgm3[k] > 0.0 ? spec->layer_density_kgm3[k]
                                                       : astraea_hip::material(spec->layer_material[k]).rho0;
        }
        fEnv = f[nl - 1];
        if (!(fEnv > 0.0 && fEnv < 1.0)) return false;
        const double rPure = kPi * std::sqrt(K / (2.0 * kPi * kGsi));
        dr = rPure / 20000.0;
        return true;
    }
    double coreTop(double M, double* radii) const {
        double v = 0.0;
        for (int k = 0; k < nc; ++k) {
            v += f[k] * M / rho[k];
            if (radii) radii[k] = std::cbrt(3.0 * v / (4.0 * kPi));
        }
        return std::cbrt(3.0 * v / (4.0 * kPi));
    }
    // density at the core's top that puts the envelope's mass fraction inside its surface
    double solveTop(double M) const {
        const double rc = coreTop(M, nullptr), mc = (1.0 - fEnv) * M;
        double lo = std::log(1.0e-8), hi = std::log(1.0e8);
        for (int it = 0; it < 90; ++it) {
            const double mid = 0.5 * (lo + hi);
            if (integrateEnvelope(K, rc, mc, std::exp(mid), dr, false).mEnvelope < fEnv * M) lo = mid; else hi = mid;
        }
        return std::exp(0.5 * (lo + hi));
    }
    // surface radius of total mass M (it grows with the mass: the core pushes the envelope out)
    double surfaceOf(double M) const {
        return integrateEnvelope(K, coreTop(M, nullptr), (1.0 - fEnv) * M, solveTop(M), dr, false).rSurface;
    }
};

int buildGasGiant(const astraea_body_spec* spec, uint32_t n, astraea_body_out* out, double* outMass, double* outSpacing) {
    GasGiantModel G;
    if (!G.init(spec)) return ASTRAEA_E_INVALID;
    const int nc = G.nc;
    const double K = G.K, R = spec->radius_m, fEnv = G.fEnv, dr = G.dr;
    const double* rho = G.rho;
    auto coreTop = [&](double M, double* radii) { return G.coreTop(M, radii); };
    auto solveTop = [&](double M) { return G.solveTop(M); };
    auto surfaceOf = [&](double M) { return G.surfaceOf(M); };
    double lo = std::log(1.0e18), hi = std::log(1.0e31);
    if (!(surfaceOf(std::exp(lo)) < R && surfaceOf(std::exp(hi)) > R)) return ASTRAEA_E_INVALID;
    for (int it = 0; it < 70; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (surfaceOf(std::exp(mid)) < R) lo = mid; else hi = mid;
    }
    const double M = std::exp(0.5 * (lo + hi));
    double rCore[ASTRAEA_MAX_LAYERS];
    const double rc = coreTop(M, rCore);
    const double rhoTop = solveTop(M);
    const Envelope env = integrateEnvelope(K, rc, (1.0 - fEnv) * M, rhoTop, dr, true);

    // core layers: uniform density, pressure from the envelope's base pressure K rho_top^2 inward
    std::vector<Layer> L(nc);
    double mEnc = 0.0, rPrev = 0.0;
    for (int k = 0; k < nc; ++k) {
        L[k] = {rPrev, rCore[k], rho[k], mEnc, spec->layer_material[k]};
        mEnc += 4.0 / 3.0 * kPi * rho[k] * (rCore[k] * rCore[k] * rCore[k] - rPrev * rPrev * rPrev);
        rPrev = rCore[k];
    }
    std::vector<double> pTop(nc, 0.0);
    pTop[nc - 1] = K * rhoTop * rhoTop;
    for (int k = nc - 2; k >= 0; --k) {
        const Layer& l = L[k + 1];
        const double kk = 4.0 / 3.0 * kPi * l.rho;
        const double c = l.mEncIn - kk * l.rIn * l.rIn * l.rIn;
        pTop[k] = pTop[k + 1] + kGsi * l.rho * (c * (1.0 / l.rIn - 1.0 / l.rOut) + 0.5 * kk * (l.rOut * l.rOut - l.rIn * l.rIn));
    }

    std::vector<LatticePoint> pts;
    double s0 = 0.0;
    if (!latticeBall(n, 1.0, spec->jitter, spec->seed, pts, s0)) return ASTRAEA_E_INVALID;
    // rank in radius -> enclosed mass -> new radius
    std::vector<uint32_t> order(n);
    for (uint32_t i = 0; i < n; ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        return pts[a].r2 < pts[b].r2 || (pts[a].r2 == pts[b].r2 && pts[a].id < pts[b].id);
    });
    const double mP = M / (double)n;
    const double mCoreTot = (1.0 - fEnv) * M;
    std::vector<double> rNew(n);
    for (uint32_t rank = 0; rank < n; ++rank) {
        const double mt = ((double)rank + 0.5) * mP;
        double r;
        if (mt <= mCoreTot) {
            // invert the piecewise-uniform core
            int k = 0;
            while (k + 1 < nc && mt > L[k + 1].mEncIn) ++k;
            const double a = L[k].rIn;
            r = std::cbrt(a * a * a + (mt - L[k].mEncIn) / (4.0 / 3.0 * kPi * L[k].rho));
        } else {
            const auto it = std::lower_bound(env.m.begin(), env.m.end(), mt);
            const size_t j = (size_t)std::min<std::ptrdiff_t>(std::max<std::ptrdiff_t>(it - env.m.begin(), 1),
                                                              (std::ptrdiff_t)env.m.size() - 1);
            const double t = (mt - env.m[j - 1]) / std::max(env.m[j] - env.m[j - 1], 1.0e-300);
            r = env.r[j - 1] + t * (env.r[j] - env.r[j - 1]);
        }
        rNew[order[rank]] = r;
    }
    const double rhoFloor = 1.0e-3 * rhoTop;
    uint64_t frng = spec->seed ^ 0x5851F42D4C957F2DULL;
    std::vector<double> px(n), py(n), pz(n);
    double cx = 0.0, cy = 0.0, cz = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        const double r0 = std::sqrt(pts[i].r2);
        const double sc = r0 > 0.0 ? rNew[i] / r0 : 0.0;
        px[i] = pts[i].x * sc; py[i] = pts[i].y * sc; pz[i] = pts[i].z * sc;
        cx += px[i]; cy += py[i]; cz += pz[i];
    }
    cx /= n; cy /= n; cz /= n;
    for (uint32_t i = 0; i < n; ++i) {
        const double r = rNew[i];
        int mat;
        double rhoSi, u;
        if (r <= rc && nc > 0) {
            int k = 0;
            while (k + 1 < nc && r > L[k].rOut) ++k;
            mat = L[k].mat;
            rhoSi = L[k].rho;
            const double tS = spec->surface_temp_k > 0.0 ? spec->surface_temp_k : astraea_hip::material(mat).tRef;
            u = energyForPressure(mat, rhoSi / astraea_hip::kRhoUnit, pressureAt(L, pTop, r), tS, 5.0 / 3.0);
        } else {
            mat = 3;
            const auto it = std::lower_bound(env.r.begin(), env.r.end(), r);
            const size_t j = (size_t)std::min<std::ptrdiff_t>(std::max<std::ptrdiff_t>(it - env.r.begin(), 1),
                                                              (std::ptrdiff_t)env.r.size() - 1);
            const double t = (r - env.r[j - 1]) / std::max(env.r[j] - env.r[j - 1], 1.0e-300);
            rhoSi = std::max(env.rho[j - 1] + t * (env.rho[j] - env.rho[j - 1]), rhoFloor);
            u = astraea_hip::coldEnergyPsu(3, rhoSi / astraea_hip::kRhoUnit);  // P = K rho^2, no thermal part
        }
        const double vCell = mP / rhoSi;
        if (out->x) out->x[i] = (float)((px[i] - cx) / kLUnit);
        if (out->y) out->y[i] = (float)((py[i] - cy) / kLUnit);
        if (out->z) out->z[i] = (float)((pz[i] - cz) / kLUnit);
        if (out->m) out->m[i] = (float)(mP / kMUnit);
        if (out->h) out->h[i] = (float)(spec->h_over_spacing * std::cbrt(std::sqrt(2.0) * vCell) / kLUnit);
        if (out->rho) out->rho[i] = (float)(rhoSi / astraea_hip::kRhoUnit);
        if (out->u) out->u[i] = (float)u;
        if (out->mat) out->mat[i] = (uint8_t)mat;
        if (out->eps_fail) {
            double e = 1.0;
            if (spec->strength_flaws && mat != 3) {
                const double U = std::max(uniform01(frng), 1.0e-300);
                e = std::pow(-std::log(U) / (spec->weibull_k * vCell), 1.0 / spec->weibull_m);
            }
            out->eps_fail[i] = (float)e;
        }
    }
    if (outMass) *outMass = M;
    if (outSpacing) *outSpacing = std::cbrt(std::sqrt(2.0) * mP / rho[0]);
    return ASTRAEA_OK;
}

}  // namespace

extern "C" {

uint64_t astraea_splitmix64(uint64_t x) { return splitmix(x); }

int astraea_body_radius_for_mass(const astraea_body_spec* spec, double massKg, double* radiusM) {
    if (!spec || !radiusM || !(massKg > 0.0) || spec->num_layers < 1 || spec->num_layers > ASTRAEA_MAX_LAYERS)
        return ASTRAEA_E_INVALID;
    const int nl = (int)spec->num_layers;
    for (int k = 0; k + 1 < nl; ++k)
        if (spec->layer_material[k] == 3) return ASTRAEA_E_INVALID;
    if (spec->layer_material[nl - 1] == 3) {
        if (nl < 2) return ASTRAEA_E_INVALID;
        GasGiantModel G;
        if (!G.init(spec)) return ASTRAEA_E_INVALID;
        *radiusM = G.surfaceOf(massKg);
        return std::isfinite(*radiusM) && *radiusM > 0.0 ? ASTRAEA_OK : ASTRAEA_E_INVALID;
    }
    double fsum = 0.0, vsum = 0.0;
    for (int k = 0; k < nl; ++k) {
        const double rho = spec->layer_density_kgm3[k] > 0.0 ? spec->layer_density_kgm3[k]
                                                             : astraea_hip::material(spec->layer_material[k]).rho0;
        fsum += spec->layer_mass_frac[k];
        vsum += spec->layer_mass_frac[k] / rho;
    }
    if (!(fsum > 0.0)) return ASTRAEA_E_INVALID;
    *radiusM = std::cbrt(3.0 * massKg * (vsum / fsum) / (4.0 * kPi));
    return ASTRAEA_OK;
}

void astraea_default_body_spec(astraea_body_spec* s) {
    if (!s) return;
    std::memset(s, 0, sizeof(*s));
    s->radius_m = 6.371e6;
    s->num_layers = 2;
    s->layer_material[0] = 0;  // iron core
    s->layer_material[1] = 1;  // granite mantle
    s->layer_mass_frac[0] = 0.3;
    s->layer_mass_frac[1] = 0.7;
    s->layer_density_kgm3[0] = 0.0;
    s->layer_density_kgm3[1] = 0.0;
    s->h_over_spacing = 1.3;
    s->jitter = 0.1;
    s->surface_temp_k = 0.0;
    s->strength_flaws = 0;
    s->weibull_m = 9.0;
    s->weibull_k = 4.0e29;
    s->seed = 42;
}

int astraea_build_body(const astraea_body_spec* spec, uint32_t n, astraea_body_out* out, double* outMass,
                       double* outSpacing) {
    if (!spec || !out || n == 0 || spec->num_layers < 1 || spec->num_layers > ASTRAEA_MAX_LAYERS ||
        !(spec->radius_m > 0.0))
        return ASTRAEA_E_INVALID;
    const int nl = (int)spec->num_layers;
    for (int k = 0; k + 1 < nl; ++k)
        if (spec->layer_material[k] == 3) return ASTRAEA_E_INVALID;  // H/He only as the outermost layer
    if (spec->layer_material[nl - 1] == 3) return nl >= 2 ? buildGasGiant(spec, n, out, outMass, outSpacing) : ASTRAEA_E_INVALID;
    const double R = spec->radius_m;
    // Layer radii from mass fractions and densities: r_k^3 = R^3 * (sum_{j<=k} f_j/rho_j) / (sum_j f_j/rho_j)
    double fsum = 0.0, vsum = 0.0, rho[ASTRAEA_MAX_LAYERS];
    for (int k = 0; k < nl; ++k) {
        rho[k] = spec->layer_density_kgm3[k] > 0.0 ? spec->layer_density_kgm3[k]
                                                   : astraea_hip::material(spec->layer_material[k]).rho0;
        fsum += spec->layer_mass_frac[k];
        vsum += spec->layer_mass_frac[k] / rho[k];
    }
    if (!(fsum > 0.0)) return ASTRAEA_E_INVALID;
    std::vector<Layer> L(nl);
    double acc = 0.0, mEnc = 0.0, rPrev = 0.0;
    for (int k = 0; k < nl; ++k) {
        acc += spec->layer_mass_frac[k] / rho[k];
        const double rOut = (k + 1 == nl) ? R : R * std::cbrt(acc / vsum);
        L[k] = {rPrev, rOut, rho[k], mEnc, spec->layer_material[k]};
        mEnc += 4.0 / 3.0 * kPi * rho[k] * (rOut * rOut * rOut - rPrev * rPrev * rPrev);
        rPrev = rOut;
    }
    const double mTotal = mEnc;
    std::vector<double> pTop(nl, 0.0);
    for (int k = nl - 2; k >= 0; --k) {
        // pressure at the top of layer k = pressure at the bottom of layer k+1
        const Layer& l = L[k + 1];
        const double kk = 4.0 / 3.0 * kPi * l.rho;
        const double c = l.mEncIn - kk * l.rIn * l.rIn * l.rIn;
        pTop[k] = pTop[k + 1] + kGsi * l.rho * (c * (1.0 / l.rIn - 1.0 / l.rOut) + 0.5 * kk * (l.rOut * l.rOut - l.rIn * l.rIn));
    }

    // HCP lattice: number density sqrt(2)/s^3; the n points nearest the centre, centred; layers are
    // assigned by radius, then the body is shifted to its centre of mass and the hydrostatic
    // pressure is evaluated at the final radius
    std::vector<LatticePoint> pts;
    double s = 0.0;
    if (!latticeBall(n, R, spec->jitter, spec->seed, pts, s)) return ASTRAEA_E_INVALID;
    const double V = 4.0 / 3.0 * kPi * R * R * R;
    const double vCell = V / (double)n;
    const double hPsu = spec->h_over_spacing * s / kLUnit;
    std::vector<double> mass(n);
    std::vector<int> lay(n);
    double cx = 0.0, cy = 0.0, cz = 0.0, mt = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        const double r = std::sqrt(pts[i].r2);
        int k = 0;
        while (k + 1 < nl && r > L[k].rOut) ++k;
        lay[i] = k;
        mass[i] = L[k].rho * vCell;
        cx += mass[i] * pts[i].x; cy += mass[i] * pts[i].y; cz += mass[i] * pts[i].z; mt += mass[i];
    }
    cx /= mt; cy /= mt; cz /= mt;
    const double scale = mTotal / mt;  // equal-volume cells, masses rescaled to the layered total
    uint64_t frng = spec->seed ^ 0x5851F42D4C957F2DULL;
    for (uint32_t i = 0; i < n; ++i) {
        const Layer& l = L[lay[i]];
        const int mat = l.mat;
        const double px = pts[i].x - cx, py = pts[i].y - cy, pz = pts[i].z - cz;
        const double r = std::sqrt(px * px + py * py + pz * pz);
        const double rhoPsu = l.rho / astraea_hip::kRhoUnit;
        const double tS = spec->surface_temp_k > 0.0 ? spec->surface_temp_k : astraea_hip::material(mat).tRef;
        const double u = energyForPressure(mat, rhoPsu, pressureAt(L, pTop, r), tS, 1.4);
        if (out->x) out->x[i] = (float)(px / kLUnit);
        if (out->y) out->y[i] = (float)(py / kLUnit);
        if (out->z) out->z[i] = (float)(pz / kLUnit);
        if (out->m) out->m[i] = (float)(mass[i] * scale / kMUnit);
        if (out->h) out->h[i] = (float)hPsu;
        if (out->rho) out->rho[i] = (float)rhoPsu;
        if (out->u) out->u[i] = (float)u;
        if (out->mat) out->mat[i] = (uint8_t)mat;
        if (out->eps_fail) {
            double e = 1.0;
            if (spec->strength_flaws) {
                // Weibull flaw threshold of one particle: k V eps^m = -ln U (Benz & Asphaug 1995).
                const double U = std::max(uniform01(frng), 1.0e-300);
                e = std::pow(-std::log(U) / (spec->weibull_k * vCell), 1.0 / spec->weibull_m);
            }
            out->eps_fail[i] = (float)e;
        }
    }
    if (outMass) *outMass = mTotal;
    if (outSpacing) *outSpacing = s;
    return ASTRAEA_OK;
}

int astraea_place_two_bodies(uint32_t nt, uint32_t ni, float* x, float* y, float* z, float* vx, float* vy, float* vz,
                             const float* m, double rT, double rI, double vRatio, double bImp, double sep,
                             double* outVesc) {
    if (!x || !y || !z || !vx || !vy || !vz || !m || nt == 0 || ni == 0) return ASTRAEA_E_INVALID;
    const double G = 66.7430;  // PSU, engine/units.hpp:26
    double mT = 0.0, mI = 0.0;
    for (uint32_t i = 0; i < nt; ++i) mT += (double)m[i];
    for (uint32_t i = 0; i < ni; ++i) mI += (double)m[nt + i];
    const double mTot = mT + mI;
    const double rc = std::max(rT + rI, 1.0e-3);
    const double vEsc = std::sqrt(2.0 * G * mTot / rc);
    const double vr = std::max(0.0, vRatio);
    const double vc = vr * vEsc;
    double d0 = std::max(sep, 1.0) * rc;
    double v0sq = vc * vc - 2.0 * G * mTot * (1.0 / rc - 1.0 / d0);
    if (v0sq < 0.0) {
        const double rApo = rc / std::max(1.0 - vr * vr, 1.0e-6);
        d0 = std::min(std::max(0.98 * rApo, 1.01 * rc), d0);
        v0sq = std::max(0.0, vc * vc - 2.0 * G * mTot * (1.0 / rc - 1.0 / d0));
    }
    const double v0 = std::sqrt(v0sq);
    const double b = std::min(std::max(bImp, -1.2), 1.2);
    double vPerp = rc * vc * b / d0;
    if (std::abs(vPerp) > 0.995 * v0) vPerp = (vPerp >= 0.0 ? 1.0 : -1.0) * 0.995 * v0;
    const double vRad = -std::sqrt(std::max(0.0, v0 * v0 - vPerp * vPerp));
    const double fT = -mI / mTot, fI = mT / mTot;
    for (uint32_t i = 0; i < nt + ni; ++i) {
        const double f = i < nt ? fT : fI;
        x[i] = (float)((double)x[i] + f * d0);
        vx[i] = (float)((double)vx[i] + f * vRad);
        vz[i] = (float)((double)vz[i] + f * (-vPerp));
    }
    // zero system COM and momentum
    double c[3] = {0, 0, 0}, p[3] = {0, 0, 0}, mt = 0.0;
    for (uint32_t i = 0; i < nt + ni; ++i) {
        const double mm = (double)m[i];
        mt += mm;
        c[0] += mm * x[i]; c[1] += mm * y[i]; c[2] += mm * z[i];
        p[0] += mm * vx[i]; p[1] += mm * vy[i]; p[2] += mm * vz[i];
    }
    for (int a = 0; a < 3; ++a) { c[a] /= mt; p[a] /= mt; }
    for (uint32_t i = 0; i < nt + ni; ++i) {
        x[i] = (float)((double)x[i] - c[0]); y[i] = (float)((double)y[i] - c[1]); z[i] = (float)((double)z[i] - c[2]);
        vx[i] = (float)((double)vx[i] - p[0]); vy[i] = (float)((double)vy[i] - p[1]); vz[i] = (float)((double)vz[i] - p[2]);
    }
    if (outVesc) *outVesc = vEsc;
    return ASTRAEA_OK;
}

}  // extern "C"
