// Host geometry -> flag fields for the FoilCFD solver core: NACA 4-digit
// sections, domain boundary flags, scanline-parity airfoil voxelization with
// z extrusion, trailing-edge gap closure and the refinement-patch masks.
// Behaviour (and the FP32 expression order, so flag fields come out
// bit-identical under the same host compiler with -ffp-contract=off) follows
// the reference's src/geom/airfoil.cpp and src/geom/voxelizer.cpp.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "fc_layout.hpp"

namespace fc {

struct Vec2 {
    float x = 0.0f, y = 0.0f;
};

struct Airfoil {
    std::string name;
    std::vector<Vec2> points;  // closed Selig loop TE -> upper -> LE -> lower -> TE
    int leIndex = 0;
    bool valid() const { return points.size() >= 20; }
};

// Where the foil sits in the lattice (reference DomainLayout, voxelizer.h:57-85).
struct DomainLayout {
    Dims dims;
    int chordCells = 256;
    float anchorXFrac = 0.3f;
    float anchorYFrac = 0.5f;
    float anchorXCells = -1.0f;
    float anchorYCells = -1.0f;
    bool anchorAbsolute = false;
    float anchorX() const {
        return (anchorAbsolute || anchorXCells >= 0.0f) ? anchorXCells
                                                        : anchorXFrac * static_cast<float>(dims.nx);
    }
    float anchorY() const {
        return (anchorAbsolute || anchorYCells >= 0.0f) ? anchorYCells
                                                        : anchorYFrac * static_cast<float>(dims.ny);
    }
};

// Refinement patch in coarse cells, [x0,x1) x [y0,y1) x full z
// (reference PatchBox, lbm_refine.cuh:54-61).
struct PatchBox {
    int x0 = 0, y0 = 0, x1 = 0, y1 = 0;
    bool valid() const { return (x1 - x0) > 8 && (y1 - y0) > 8; }
};

inline constexpr int kInterfaceShellFine = 2;   // lbm_refine.cuh:71
inline constexpr int kRestrictBandCoarse = 2;   // lbm_refine.cuh:77
inline constexpr int kRestrictBlendCoarse = 3;  // lbm_refine.cuh:85

// NACA 4-digit section, closed trailing edge, cosine spacing; empty points and
// a reason string on bad input (reference generateNACA4, airfoil.cpp:177-266).
bool generateNACA4(const std::string& digits, int pointsPerSurface, Airfoil& out,
                   std::string* reason);

std::vector<std::uint8_t> buildBoundaryFlags(const Dims& d);
void voxelizeAirfoil(const Airfoil& foil, float aoaDeg, const DomainLayout& layout,
                     std::vector<std::uint8_t>& flags);
int closeTrailingEdgeGaps(const Dims& d, std::vector<std::uint8_t>& flags);
std::vector<std::uint8_t> buildCleanFoilFlags(const Airfoil& foil, float aoaDeg,
                                              const DomainLayout& layout);

Dims fineDimsFor(const PatchBox& box, const Dims& coarse, int factor);
DomainLayout makeFineLayout(const DomainLayout& coarse, const PatchBox& box, int factor);
void stampInterfaceShell(const Dims& d, std::vector<std::uint8_t>& flags);
std::vector<std::uint8_t> buildFinePatchFlags(const Airfoil& foil, float aoaDeg,
                                              const DomainLayout& coarse, const PatchBox& box,
                                              int factor);
PatchBox derivePatchBox(const Dims& d, const std::vector<std::uint8_t>& flags, int marginX0,
                        int marginX1, int marginY0, int marginY1, int clearance);

}  // namespace fc
