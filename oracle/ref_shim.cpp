// FoilCFD oracle -- TEST INFRASTRUCTURE ONLY.
//
// C entry points over the parts of the reference that compile from their own
// sources with the host compiler alone: the header-only unit conversions
// (src/sim/units.h) and the NACA 4-digit generator (src/geom/airfoil.cpp,
// linked with src/platform/platform.cpp). Built by oracle/Makefile into
// oracle/_ref/libfcref.so straight from the sources under /root/reference;
// used only to write and check golden vectors, never by the product.
#include <cstring>
#include <string>

#include "geom/airfoil.h"
#include "sim/units.h"

using namespace foilcfd;

namespace {
void pack(const LatticeScaling& s, float* out) {
    out[0] = s.phys.chord_m;
    out[1] = s.phys.airspeed_ms;
    out[2] = s.phys.nu_m2s;
    out[3] = static_cast<float>(s.chordCells);
    out[4] = s.u_lat;
    out[5] = s.dx;
    out[6] = s.dt;
    out[7] = s.nu_lat_target;
    out[8] = s.nu_lat;
    out[9] = s.tau;
    out[10] = s.tauClamped ? 1.0f : 0.0f;
    out[11] = s.reTarget;
    out[12] = s.reEffective;
}
LatticeScaling scaling(float c, float u, float nu, int nc, float ulat) {
    PhysicalParams p;
    p.chord_m = c;
    p.airspeed_ms = u;
    p.nu_m2s = nu;
    return computeScaling(p, nc, ulat);
}
}  // namespace

extern "C" {

void ref_compute_scaling(float chord, float speed, float nu, int nc, float ulat, float* out13) {
    pack(scaling(chord, speed, nu, nc, ulat), out13);
}

void ref_refined_scaling(float chord, float speed, float nu, int nc, float ulat, int factor,
                         float* out13) {
    pack(refinedScaling(scaling(chord, speed, nu, nc, ulat), factor), out13);
}

float ref_ramped_tau(float chord, float speed, float nu, int nc, float ulat, long long step, int nx) {
    return rampedTau(scaling(chord, speed, nu, nc, ulat), step, nx);
}

float ref_ramped_u(float chord, float speed, float nu, int nc, float ulat, long long step, int nx) {
    return rampedU(scaling(chord, speed, nu, nc, ulat), step, nx);
}

float ref_max_effective_re(int nc, float ulat) { return maxEffectiveRe(nc, ulat); }

float ref_force_to_coefficient(float chord, float speed, float nu, int nc, float ulat, float force,
                               int span) {
    return scaling(chord, speed, nu, nc, ulat).forceToCoefficient(force, span);
}

int ref_generate_naca4(const char* digits, int pointsPerSurface, float* xy, int cap, int* leIndex) {
    const AirfoilLoadResult r = generateNACA4(digits, pointsPerSurface);
    if (!r.ok) return -1;
    const int n = static_cast<int>(r.airfoil.points.size());
    if (n > cap) return -2;
    for (int i = 0; i < n; ++i) {
        xy[2 * i] = r.airfoil.points[i].x;
        xy[2 * i + 1] = r.airfoil.points[i].y;
    }
    if (leIndex) *leIndex = r.airfoil.leIndex;
    return n;
}

}  // extern "C"
