// Physical <-> lattice unit conversion for the FoilCFD solver core (host).
//
// Same contract and the same FP32 formulas as the reference's single scaling
// authority (src/sim/units.h): dx = c/Nc, dt = u_lat dx/U, nu_lat =
// nu dt/dx^2, tau = 3 nu_lat + 1/2 clamped at 0.5005, the refined-level
// acoustic scaling, the startup viscosity / inlet-speed ramps and the force
// coefficient. The expressions keep the reference's operation order so the
// results are bit-identical under the same host compiler and -ffp-contract=off.
#pragma once

#include <algorithm>
#include <cmath>

namespace fc {

inline constexpr float kDefaultULat = 0.08f;     // units.h:20
inline constexpr float kMaxULat = 0.12f;         // units.h:24
inline constexpr float kMinTau = 0.5005f;        // units.h:29
inline constexpr float kNuAir = 1.48e-5f;        // units.h:32
inline constexpr float kTRTMagicLambda = 3.0f / 16.0f;  // units.h:37
inline constexpr float kSmagorinskyCs = 0.1733f;  // units.h:46
inline constexpr float kStartupNuMultiplier = 4.0f;  // units.h:263
inline constexpr int kStartupRampNxFactor = 2;       // units.h:266
inline constexpr float kVelocityRampFraction = 0.5f;  // units.h:300

struct Scaling {
    float chord_m = 1.2f, airspeed_ms = 30.0f, nu_m2s = kNuAir;
    int chordCells = 256;
    float u_lat = kDefaultULat;
    float dx = 0, dt = 0;
    float nu_lat_target = 0, nu_lat = 0, tau = 0;
    bool tauClamped = false;
    float reTarget = 0, reEffective = 0;

    float flowThroughSteps(int nx) const { return static_cast<float>(nx) / std::max(u_lat, 1e-6f); }
    float forceToCoefficient(float forceLat, int spanCells) const {
        const float qref = 0.5f * u_lat * u_lat * static_cast<float>(chordCells) *
                           static_cast<float>(spanCells);
        return (qref > 0.0f) ? forceLat / qref : 0.0f;
    }
};

// units.h:192-219
inline Scaling computeScaling(float chord_m, float airspeed_ms, float nu_m2s, int chordCells,
                              float u_lat = kDefaultULat) {
    Scaling s;
    s.chord_m = chord_m;
    s.airspeed_ms = airspeed_ms;
    s.nu_m2s = nu_m2s;
    s.chordCells = chordCells;
    s.u_lat = std::min(u_lat, kMaxULat);
    s.dx = chord_m / static_cast<float>(chordCells);
    s.dt = s.u_lat * s.dx / airspeed_ms;
    s.nu_lat_target = nu_m2s * s.dt / (s.dx * s.dx);
    const float tauTarget = 3.0f * s.nu_lat_target + 0.5f;
    s.tau = std::max(tauTarget, kMinTau);
    s.tauClamped = tauTarget < kMinTau;
    s.nu_lat = (s.tau - 0.5f) / 3.0f;
    s.reTarget = airspeed_ms * chord_m / nu_m2s;
    s.reEffective = s.u_lat * static_cast<float>(chordCells) / s.nu_lat;
    return s;
}

// units.h:224-227
inline float maxEffectiveRe(int chordCells, float u_lat = kDefaultULat) {
    const float nuMin = (kMinTau - 0.5f) / 3.0f;
    return std::min(u_lat, kMaxULat) * static_cast<float>(chordCells) / nuMin;
}

// units.h:242-254 (acoustic scaling of a refined level).
inline Scaling refinedScaling(const Scaling& parent, int factor) {
    Scaling f = parent;
    const float m = static_cast<float>(factor);
    f.chordCells = parent.chordCells * factor;
    f.dx = parent.dx / m;
    f.dt = parent.dt / m;
    f.nu_lat_target = parent.nu_lat_target * m;
    f.nu_lat = parent.nu_lat * m;
    f.tau = 3.0f * f.nu_lat + 0.5f;
    return f;
}

// Fine-level relaxation time tau_f = m (tau_c - 1/2) + 1/2 (lbm_refine.cuh:99-101).
inline float fineTauFor(float tauC, int factor) {
    return static_cast<float>(factor) * (tauC - 0.5f) + 0.5f;
}

// units.h:274-283
inline float rampedTau(const Scaling& s, long long step, int nx) {
    const long long rampSteps = static_cast<long long>(kStartupRampNxFactor) * static_cast<long long>(nx);
    if (step >= rampSteps) return s.tau;
    const float t = static_cast<float>(step) / static_cast<float>(rampSteps);
    const float nuNow = s.nu_lat * (kStartupNuMultiplier + t * (1.0f - kStartupNuMultiplier));
    return 3.0f * nuNow + 0.5f;
}

// units.h:310-318
inline float rampedU(const Scaling& s, long long step, int nx) {
    const long long rampSteps = static_cast<long long>(
        kVelocityRampFraction * static_cast<float>(kStartupRampNxFactor) * static_cast<float>(nx));
    if (rampSteps <= 0 || step >= rampSteps) return s.u_lat;
    const float t = static_cast<float>(step) / static_cast<float>(rampSteps);
    const float sm = t * t * (3.0f - 2.0f * t);
    return s.u_lat * sm;
}

// Smagorinsky prefactor 18 sqrt(2) Cs^2 as the reference launch wrapper forms
// it (src/sim/lbm_core.cu:785-787).
inline float smagPrefactor(float cs) {
    return (cs > 0.0f) ? 18.0f * 1.41421356f * cs * cs : 0.0f;
}

}  // namespace fc
