// VG vortex-strength audit math on the host (reference src/geom/vg_audit.h,
// vg_audit.cpp:12-101): the Wendt correlation for the circulation a vane
// sheds, and the signed streamwise-vorticity integrals over a band of a
// crossflow plane held in host memory. The live-field variant of the band
// integral runs on the device (fc_bl.hip).
#include <algorithm>
#include <cmath>

#include "../../include/foilcfd.h"

namespace {

// Wendt 2001 regression constants as used by Dudek 2005 Eq. 3 (vg_audit.h:27-29).
constexpr float kK1 = 1.61f;
constexpr float kK2 = 0.48f;
constexpr float kK3 = 1.41f;
constexpr float kPi = 3.14159265f;

inline float at(const float* a, int ny, int y, int z) {
    return a[static_cast<std::size_t>(y) + static_cast<std::size_t>(ny) * static_cast<std::size_t>(z)];
}

// omega_x = dw/dy - dv/dz at (y, z): central differences inside the band,
// the in-band one-sided y stencil on its first and last row, z periodic.
inline float omegaX(const float* v, const float* w, int ny, int nz, int y, int z, int yMin, int yMax) {
    const int zm = z == 0 ? nz - 1 : z - 1;
    const int zp = z == nz - 1 ? 0 : z + 1;
    float dwdy;
    if (y == yMin)
        dwdy = at(w, ny, y + 1, z) - at(w, ny, y, z);
    else if (y == yMax)
        dwdy = at(w, ny, y, z) - at(w, ny, y - 1, z);
    else
        dwdy = 0.5f * (at(w, ny, y + 1, z) - at(w, ny, y - 1, z));
    const float dvdz = 0.5f * (at(v, ny, y, zp) - at(v, ny, y, zm));
    return dwdy - dvdz;
}

inline void addSigned(fc_plane_circulation& c, float om) {
    if (om > 0.0f)
        c.pos += om;
    else
        c.neg += om;
}

}  // namespace

extern "C" {

float fc_wendt_circulation_lattice(float beta_rad, float ue_lat, float h_cells, float vane_chord_cells,
                                   float delta99_cells) {
    if (ue_lat <= 0.0f || h_cells <= 0.0f || vane_chord_cells <= 0.0f || delta99_cells <= 0.0f) return 0.0f;
    // wall-mounted half wing: mirror doubles the span, 4/pi elliptic equivalence
    const float aspect = 8.0f * h_cells / (kPi * vane_chord_cells);
    const float bl = std::tanh(kK3 * h_cells / delta99_cells);
    return kK1 * std::fabs(beta_rad) * ue_lat * vane_chord_cells * bl / (1.0f + kK2 / aspect);
}

int fc_integrate_streamwise_circulation(const float* v, const float* w, int ny, int nz, int y_min, int y_max,
                                        fc_plane_circulation* out) {
    if (!out) return FC_ERR_INVALID;
    *out = fc_plane_circulation{0.0f, 0.0f};
    if (!v || !w || ny <= 0 || nz < 3) return FC_OK;
    y_min = std::max(y_min, 0);
    y_max = std::min(y_max, ny - 1);
    if (y_max <= y_min) return FC_OK;
    for (int z = 0; z < nz; ++z)
        for (int y = y_min; y <= y_max; ++y) addSigned(*out, omegaX(v, w, ny, nz, y, z, y_min, y_max));
    return FC_OK;
}

int fc_integrate_band_circulation(const float* v, const float* w, int ny, int nz, int y_min, int y_max, int z0,
                                  int z1, fc_band_circulation* out) {
    if (!out) return FC_ERR_INVALID;
    *out = fc_band_circulation{};
    if (!v || !w || ny <= 0 || nz < 3) return FC_OK;
    y_min = std::max(y_min, 0);
    y_max = std::min(y_max, ny - 1);
    z0 = std::max(z0, 0);
    z1 = std::min(z1, nz - 1);
    if (y_max <= y_min || z1 < z0) return FC_OK;
    for (int z = 0; z < nz; ++z) {
        const bool zone = z >= z0 && z <= z1;
        fc_plane_circulation& dst = zone ? out->zone : out->ambient;
        ++(zone ? out->zone_cols : out->ambient_cols);
        for (int y = y_min; y <= y_max; ++y) addSigned(dst, omegaX(v, w, ny, nz, y, z, y_min, y_max));
    }
    return FC_OK;
}

}  // extern "C"
