// Host construction of the wall-model cell list (fluid cells whose pull
// links touch the foil), with occupancy-gradient wall normals, velocity
// sampling cells and bounce-back link masks. Same selection, normals,
// sampling fallback and ordering as the reference (src/sim/wall_model.cpp),
// so lists are bit-identical for the same flag fields.
#pragma once

#include <cstdint>
#include <vector>

#include "fc_layout.hpp"

namespace fc {

inline constexpr int kWallNormalRadius = 2;       // wall_model.h:29
inline constexpr int kWallSampleCellsCoarse = 2;  // wall_model.h:36
inline constexpr int kWallSampleCellsFine = 3;    // wall_model.h:37

struct WallList {
    std::vector<long long> cellIdx, sampleIdx;
    std::vector<float> nx, ny, nz, sampleDist;
    std::vector<std::uint32_t> linkMask;
    std::size_t size() const { return cellIdx.size(); }
};

struct WallListStats {
    int listed = 0, excludedVG = 0, degenerate = 0;
};

// active: live flags (foil + vanes); clean: vane-free flags (same size).
WallList buildWallCellList(const Dims& d, const std::vector<std::uint8_t>& active,
                           const std::vector<std::uint8_t>& clean, int sampleCells,
                           WallListStats* stats);

}  // namespace fc
