// Host geometry for the FoilCFD solver core (see fc_geometry.hpp).
// Compiled with -ffp-contract=off: the flag fields must be bit-identical to
// the reference's host voxelizer, which is ordinary FP32 host arithmetic.
#include "fc_geometry.hpp"

#include <algorithm>
#include <cmath>
#include <limits>

namespace fc {
namespace {

constexpr float kPi = 3.14159265358979f;  // the constant the reference uses

inline std::size_t at(const Dims& d, int x, int y, int z) {
    return static_cast<std::size_t>(x) +
           static_cast<std::size_t>(d.nx) *
               (static_cast<std::size_t>(y) + static_cast<std::size_t>(d.ny) * static_cast<std::size_t>(z));
}

}  // namespace

bool generateNACA4(const std::string& digits, int pointsPerSurface, Airfoil& out,
                   std::string* reason) {
    out = Airfoil{};
    auto fail = [&](const std::string& why) {
        if (reason) *reason = why;
        return false;
    };
    if (digits.size() != 4) return fail("NACA 4-digit designation must have 4 characters: \"" + digits + "\"");
    for (char ch : digits)
        if (ch < '0' || ch > '9') return fail("NACA designation \"" + digits + "\" is not all decimal digits");
    // m = max camber, p = its chordwise position, t = max thickness.
    const float m = static_cast<float>(digits[0] - '0') / 100.0f;
    const float p = static_cast<float>(digits[1] - '0') / 10.0f;
    const float t = static_cast<float>((digits[2] - '0') * 10 + (digits[3] - '0')) / 100.0f;
    if (t <= 0.0f) return fail("NACA thickness of 0 % chord has no section");
    if (m > 0.0f && p <= 0.0f) return fail("cambered NACA section needs a camber position > 0");

    const int n = std::max(pointsPerSurface, 20);
    out.name = "NACA " + digits;
    out.points.reserve(static_cast<std::size_t>(2 * n - 1));

    // Station k of n: cosine-spaced x, closed-TE thickness (last coefficient
    // -0.1036 so yt(1) = 0), two-arc camber line, thickness applied normal
    // to the camber line.
    auto station = [&](int k, Vec2& up, Vec2& lo) {
        const float x = 0.5f * (1.0f - std::cos(kPi * static_cast<float>(k) / static_cast<float>(n - 1)));
        const float sx = std::sqrt(x);
        const float yt = 5.0f * t *
                         (0.2969f * sx - 0.1260f * x - 0.3516f * x * x + 0.2843f * x * x * x -
                          0.1036f * x * x * x * x);
        float yc = 0.0f, slope = 0.0f;
        if (m > 0.0f) {
            if (x < p) {
                yc = (m / (p * p)) * (2.0f * p * x - x * x);
                slope = (2.0f * m / (p * p)) * (p - x);
            } else {
                yc = (m / ((1.0f - p) * (1.0f - p))) * ((1.0f - 2.0f * p) + 2.0f * p * x - x * x);
                slope = (2.0f * m / ((1.0f - p) * (1.0f - p))) * (p - x);
            }
        }
        const float th = std::atan(slope);
        const float st = std::sin(th);
        const float ct = std::cos(th);
        up = Vec2{x - yt * st, yc + yt * ct};
        lo = Vec2{x + yt * st, yc - yt * ct};
    };
    Vec2 u, l;
    for (int k = n - 1; k >= 0; --k) {  // TE -> LE along the upper surface
        station(k, u, l);
        out.points.push_back(u);
    }
    out.leIndex = n - 1;
    for (int k = 1; k < n; ++k) {  // LE (shared) -> TE along the lower surface
        station(k, u, l);
        out.points.push_back(l);
    }
    return true;
}

std::vector<std::uint8_t> buildBoundaryFlags(const Dims& d) {
    std::vector<std::uint8_t> flags(static_cast<std::size_t>(d.cells()), kFluid);
    for (int z = 0; z < d.nz; ++z) {
        for (int y = 0; y < d.ny; ++y) {
            const std::size_t row = at(d, 0, y, z);
            // Slip planes first; the inlet then wins its corners, while the
            // outlet yields to the slip planes (reference voxelizer.cpp:28-60).
            if (y == 0)
                for (int x = 0; x < d.nx; ++x) flags[row + x] = kSlipBottom;
            if (y == d.ny - 1)
                for (int x = 0; x < d.nx; ++x) flags[row + x] = kSlipTop;
            flags[row] = kInlet;
            if (y != 0 && y != d.ny - 1) flags[row + d.nx - 1] = kOutlet;
        }
    }
    return flags;
}

void voxelizeAirfoil(const Airfoil& foil, float aoaDeg, const DomainLayout& layout,
                     std::vector<std::uint8_t>& flags) {
    if (!foil.valid()) return;
    const Dims& d = layout.dims;
    if (d.nx < 3 || d.ny < 3 || d.nz < 1) return;

    // Rotate about the quarter chord by -AoA (nose up for AoA > 0), scale to
    // cells, translate to the anchor.
    const float rad = aoaDeg * kPi / 180.0f;
    const float chord = static_cast<float>(layout.chordCells);
    const float ax = layout.anchorX();
    const float ay = layout.anchorY();
    const float cr = std::cos(-rad);
    const float sr = std::sin(-rad);
    const std::size_t n = foil.points.size();
    std::vector<Vec2> poly(n);
    float minY = std::numeric_limits<float>::max();
    float maxY = std::numeric_limits<float>::lowest();
    for (std::size_t i = 0; i < n; ++i) {
        const float px = foil.points[i].x - 0.25f;
        const float py = foil.points[i].y - 0.0f;
        const float qx = px * cr - py * sr;
        const float qy = px * sr + py * cr;
        poly[i] = Vec2{ax + qx * chord, ay + qy * chord};
        minY = std::min(minY, poly[i].y);
        maxY = std::max(maxY, poly[i].y);
    }

    // Scanline parity per row at the cell-centre line y + 0.5, half-open edge
    // rule so each vertex counts once; cells whose centre x + 0.5 lies
    // strictly between a crossing pair are inside.
    const int y0 = std::clamp(static_cast<int>(std::floor(minY - 0.5f)), 1, d.ny - 2);
    const int y1 = std::clamp(static_cast<int>(std::ceil(maxY - 0.5f)), 1, d.ny - 2);
    std::vector<std::uint8_t> mask(static_cast<std::size_t>(d.nx) * static_cast<std::size_t>(d.ny), 0);
    std::vector<float> xs;
    xs.reserve(8);
    for (int y = y0; y <= y1; ++y) {
        const float yc = static_cast<float>(y) + 0.5f;
        xs.clear();
        for (std::size_t i = 0; i < n; ++i) {
            const Vec2& a = poly[i];
            const Vec2& b = poly[(i + 1) % n];
            if ((a.y > yc) != (b.y > yc)) {
                const float tt = (yc - a.y) / (b.y - a.y);
                xs.push_back(a.x + tt * (b.x - a.x));
            }
        }
        std::sort(xs.begin(), xs.end());
        const std::size_t rowBase = static_cast<std::size_t>(y) * static_cast<std::size_t>(d.nx);
        for (std::size_t k = 0; k + 1 < xs.size(); k += 2) {
            const int xa = std::max(1, static_cast<int>(std::ceil(xs[k] - 0.5f)));
            const int xb = std::min(d.nx - 2, static_cast<int>(std::floor(xs[k + 1] - 0.5f)));
            for (int x = xa; x <= xb; ++x) mask[rowBase + x] = 1;
        }
    }
    // Extrude across z; only Fluid cells turn Solid.
    const std::size_t slice = static_cast<std::size_t>(d.nx) * static_cast<std::size_t>(d.ny);
    for (int z = 0; z < d.nz; ++z) {
        const std::size_t zb = slice * static_cast<std::size_t>(z);
        for (int y = y0; y <= y1; ++y) {
            const std::size_t rb = static_cast<std::size_t>(y) * static_cast<std::size_t>(d.nx);
            for (int x = 1; x < d.nx - 1; ++x)
                if (mask[rb + x] && flags[zb + rb + x] == kFluid) flags[zb + rb + x] = kSolid;
        }
    }
}

int closeTrailingEdgeGaps(const Dims& d, std::vector<std::uint8_t>& flags) {
    if (d.nx < 3 || d.ny < 3 || d.nz < 1) return 0;
    // One pass against an unmodified snapshot: only original Solid/Fluid/Solid
    // y-sandwiches close (reference voxelizer.cpp:155-183).
    const std::vector<std::uint8_t> before = flags;
    const std::size_t nx = static_cast<std::size_t>(d.nx);
    int closed = 0;
    for (int z = 0; z < d.nz; ++z)
        for (int y = 1; y < d.ny - 1; ++y)
            for (int x = 0; x < d.nx; ++x) {
                const std::size_t i = at(d, x, y, z);
                if (before[i] == kFluid && before[i - nx] == kSolid && before[i + nx] == kSolid) {
                    flags[i] = kSolid;
                    ++closed;
                }
            }
    return closed;
}

std::vector<std::uint8_t> buildCleanFoilFlags(const Airfoil& foil, float aoaDeg,
                                              const DomainLayout& layout) {
    std::vector<std::uint8_t> flags = buildBoundaryFlags(layout.dims);
    voxelizeAirfoil(foil, aoaDeg, layout, flags);
    closeTrailingEdgeGaps(layout.dims, flags);
    return flags;
}

Dims fineDimsFor(const PatchBox& box, const Dims& coarse, int factor) {
    Dims d;
    d.nx = factor * (box.x1 - box.x0);
    d.ny = factor * (box.y1 - box.y0);
    d.nz = factor * coarse.nz;
    return d;
}

DomainLayout makeFineLayout(const DomainLayout& coarse, const PatchBox& box, int factor) {
    DomainLayout fine;
    fine.dims = fineDimsFor(box, coarse.dims, factor);
    fine.chordCells = factor * coarse.chordCells;
    fine.anchorAbsolute = true;
    fine.anchorXCells = static_cast<float>(factor) * (coarse.anchorX() - static_cast<float>(box.x0));
    fine.anchorYCells = static_cast<float>(factor) * (coarse.anchorY() - static_cast<float>(box.y0));
    return fine;
}

void stampInterfaceShell(const Dims& d, std::vector<std::uint8_t>& flags) {
    const int s = kInterfaceShellFine;
    for (int z = 0; z < d.nz; ++z)
        for (int y = 0; y < d.ny; ++y) {
            const std::size_t rb = at(d, 0, y, z);
            if (y < s || y >= d.ny - s) {
                for (int x = 0; x < d.nx; ++x) flags[rb + x] = kInterface;
            } else {
                for (int x = 0; x < s; ++x) flags[rb + x] = kInterface;
                for (int x = d.nx - s; x < d.nx; ++x) flags[rb + x] = kInterface;
            }
        }
}

std::vector<std::uint8_t> buildFinePatchFlags(const Airfoil& foil, float aoaDeg,
                                              const DomainLayout& coarse, const PatchBox& box,
                                              int factor) {
    const DomainLayout fine = makeFineLayout(coarse, box, factor);
    std::vector<std::uint8_t> flags(static_cast<std::size_t>(fine.dims.cells()), kFluid);
    voxelizeAirfoil(foil, aoaDeg, fine, flags);
    closeTrailingEdgeGaps(fine.dims, flags);
    stampInterfaceShell(fine.dims, flags);
    return flags;
}

PatchBox derivePatchBox(const Dims& d, const std::vector<std::uint8_t>& flags, int marginX0,
                        int marginX1, int marginY0, int marginY1, int clearance) {
    int sx0 = d.nx, sx1 = -1, sy0 = d.ny, sy1 = -1;
    for (int z = 0; z < d.nz; ++z)
        for (int y = 0; y < d.ny; ++y)
            for (int x = 0; x < d.nx; ++x)
                if (flags[at(d, x, y, z)] == kSolid) {
                    sx0 = std::min(sx0, x);
                    sx1 = std::max(sx1, x);
                    sy0 = std::min(sy0, y);
                    sy1 = std::max(sy1, y);
                }
    PatchBox box;
    if (sx1 < sx0) return box;
    box.x0 = std::max(clearance, sx0 - marginX0);
    box.x1 = std::min(d.nx - clearance, sx1 + 1 + marginX1);
    box.y0 = std::max(clearance, sy0 - marginY0);
    box.y1 = std::min(d.ny - clearance, sy1 + 1 + marginY1);
    return box;
}

}  // namespace fc
