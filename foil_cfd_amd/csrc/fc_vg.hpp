// Vortex-generator geometry for the FoilCFD solver core (host): surface
// frames on the section, vane stamping by dense sub-cell sampling, the VG
// flag merge with trailing-edge closure, analytic vane slabs and the q-LIBB
// cut fractions along solid links. Same rules, constants and FP32 expression
// order as the reference (src/geom/vg.cpp:34-397, src/geom/airfoil.cpp:469-525).
#pragma once

#include <cstdint>
#include <vector>

#include "fc_geometry.hpp"

namespace fc {

struct Vec3 {
    float x = 0, y = 0, z = 0;
};

// VGParams, vg.h:24-36 (type: 0 single vane, 1 counter-rotating pair,
// 2 co-rotating array, 3 ramp).
struct VGParams {
    int type = 1;
    float x_c = 0.07f;
    float height_c = 0.01f;
    float length_h = 3.0f;
    float beta_deg = 16.0f;
    float pitch_c = 0.06f;
    float gap_h = 2.5f;
    int count = 4;
    bool commonFlowDown = true;
    bool enabled = true;
};

// Oriented box of one stamped vane, lattice cells (VaneSlab, voxelizer.h:47-55).
struct VaneSlab {
    Vec3 center, nAxis, dAxis, bAxis;
    float halfHeight = 0, halfLength = 0, halfThick = 0;
};

struct SurfaceFrame {
    Vec2 point, tangent, normal;
    bool valid = false;
};

inline constexpr float kQLibbMin = 0.05f;  // vg.h:120

SurfaceFrame surfaceFrameAt(const Airfoil& foil, float x_c, bool upper);
void voxelizeVG(const VGParams& vg, const Airfoil& foil, float aoaDeg, const DomainLayout& layout,
                std::vector<std::uint8_t>& flags, std::vector<VaneSlab>* slabs);
std::vector<std::uint8_t> buildFlagsWithVGs(const std::vector<VGParams>& vgs, const Airfoil& foil, float aoaDeg,
                                            const DomainLayout& layout, const std::vector<std::uint8_t>& clean,
                                            std::vector<VaneSlab>* slabs);
float rayOBBEntry(const Vec3& origin, const Vec3& dir, const VaneSlab& slab);

struct QLinkList {
    std::vector<long long> cellIdx;
    std::vector<std::uint8_t> dir, ffFluid;
    std::vector<float> q;
    std::size_t size() const { return cellIdx.size(); }
};

QLinkList buildVaneQLinks(const Dims& d, const std::vector<std::uint8_t>& active,
                          const std::vector<std::uint8_t>& clean, const std::vector<VaneSlab>& slabs);
void densifyQLinks(const QLinkList& links, long long ncells, std::vector<std::uint8_t>& qFrac,
                   std::vector<std::uint32_t>& ffMask);

}  // namespace fc
