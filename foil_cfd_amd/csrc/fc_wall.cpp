// Wall-model cell list (see fc_wall.hpp). The scan runs over z slabs on
// host threads; slabs are merged in z order, which reproduces the x-fastest
// sequential order exactly, so the list does not depend on the thread count.
#include "fc_wall.hpp"

#include <algorithm>
#include <cmath>
#include <thread>

namespace fc {
namespace {

constexpr float kDegenerateEps = 1e-3f;  // wall_model.cpp:18

struct Grid {
    const Dims& d;
    long long at(int x, int y, int z) const {
        return static_cast<long long>(x) +
               static_cast<long long>(d.nx) * (static_cast<long long>(y) + static_cast<long long>(d.ny) * z);
    }
    int wz(int z) const { return z < 0 ? z + d.nz : (z >= d.nz ? z - d.nz : z); }
    // x / y outside the grid read as not solid; z wraps.
    bool solid(const std::vector<std::uint8_t>& f, int x, int y, int z) const {
        if (x < 0 || x >= d.nx || y < 0 || y >= d.ny) return false;
        return f[at(x, y, wz(z))] == kSolid;
    }
};

void scan(const Dims& d, const std::vector<std::uint8_t>& active, const std::vector<std::uint8_t>& clean,
          int k0, int z0, int z1, WallList& out, WallListStats& st) {
    const Grid g{d};
    for (int z = z0; z < z1; ++z)
        for (int y = 0; y < d.ny; ++y)
            for (int x = 0; x < d.nx; ++x) {
                const long long c = g.at(x, y, z);
                if (active[c] != kFluid) continue;
                std::uint32_t mask = 0;
                bool vane = false;
                for (int q = 1; q < kQ; ++q) {
                    const int px = x - kCx[q], py = y - kCy[q], pz = z - kCz[q];
                    if (!g.solid(active, px, py, pz)) continue;
                    mask |= 1u << q;
                    if (!g.solid(clean, px, py, pz)) vane = true;
                }
                if (!mask) continue;
                if (vane) {
                    ++st.excludedVG;
                    continue;
                }
                // Normal = minus the first moment of the clean solid mask over
                // the (2r+1)^3 box around the cell.
                float gx = 0.0f, gy = 0.0f, gz = 0.0f;
                for (int dz = -kWallNormalRadius; dz <= kWallNormalRadius; ++dz)
                    for (int dy = -kWallNormalRadius; dy <= kWallNormalRadius; ++dy)
                        for (int dx = -kWallNormalRadius; dx <= kWallNormalRadius; ++dx)
                            if (g.solid(clean, x + dx, y + dy, z + dz)) {
                                gx += static_cast<float>(dx);
                                gy += static_cast<float>(dy);
                                gz += static_cast<float>(dz);
                            }
                const float gm = std::sqrt(gx * gx + gy * gy + gz * gz);
                if (gm < kDegenerateEps) {
                    ++st.degenerate;
                    continue;
                }
                const float nX = -gx / gm, nY = -gy / gm, nZ = -gz / gm;
                // Sampling cell: nearest fluid cell k normals out, k = k0..1,
                // else the wall cell itself.
                long long s = c;
                int sx = x, sy = y, sz = z;
                for (int k = k0; k >= 1; --k) {
                    const int cx = x + static_cast<int>(std::lround(nX * k));
                    const int cy = y + static_cast<int>(std::lround(nY * k));
                    const int cz = g.wz(z + static_cast<int>(std::lround(nZ * k)));
                    if (cx < 0 || cx >= d.nx || cy < 0 || cy >= d.ny) continue;
                    const long long cand = g.at(cx, cy, cz);
                    if (active[cand] != kFluid) continue;
                    s = cand;
                    sx = cx;
                    sy = cy;
                    sz = cz;
                    break;
                }
                int dzw = sz - z;  // minimal periodic image
                if (dzw > d.nz / 2) dzw -= d.nz;
                if (dzw < -d.nz / 2) dzw += d.nz;
                const float proj = static_cast<float>(sx - x) * nX + static_cast<float>(sy - y) * nY +
                                   static_cast<float>(dzw) * nZ;
                out.cellIdx.push_back(c);
                out.sampleIdx.push_back(s);
                out.nx.push_back(nX);
                out.ny.push_back(nY);
                out.nz.push_back(nZ);
                out.sampleDist.push_back(0.5f + std::max(proj, 0.0f));
                out.linkMask.push_back(mask);
                ++st.listed;
            }
}

template <class T>
void append(std::vector<T>& a, const std::vector<T>& b) {
    a.insert(a.end(), b.begin(), b.end());
}

}  // namespace

WallList buildWallCellList(const Dims& d, const std::vector<std::uint8_t>& active,
                           const std::vector<std::uint8_t>& clean, int sampleCells,
                           WallListStats* stats) {
    WallList merged;
    WallListStats total;
    const std::size_t n = static_cast<std::size_t>(d.cells());
    if (n == 0 || active.size() != n || clean.size() != n) {
        if (stats) *stats = total;
        return merged;
    }
    const int hw = std::max(1, static_cast<int>(std::thread::hardware_concurrency()));
    const int slabs = std::clamp(std::min(hw, d.nz), 1, 32);
    std::vector<WallList> parts(static_cast<std::size_t>(slabs));
    std::vector<WallListStats> pst(static_cast<std::size_t>(slabs));
    std::vector<std::thread> pool;
    for (int s = 0; s < slabs; ++s) {
        const int z0 = static_cast<int>(static_cast<long long>(d.nz) * s / slabs);
        const int z1 = static_cast<int>(static_cast<long long>(d.nz) * (s + 1) / slabs);
        pool.emplace_back(scan, std::cref(d), std::cref(active), std::cref(clean), sampleCells, z0, z1,
                          std::ref(parts[s]), std::ref(pst[s]));
    }
    for (auto& t : pool) t.join();
    for (int s = 0; s < slabs; ++s) {
        const WallList& p = parts[s];
        append(merged.cellIdx, p.cellIdx);
        append(merged.sampleIdx, p.sampleIdx);
        append(merged.nx, p.nx);
        append(merged.ny, p.ny);
        append(merged.nz, p.nz);
        append(merged.sampleDist, p.sampleDist);
        append(merged.linkMask, p.linkMask);
        total.listed += pst[s].listed;
        total.excludedVG += pst[s].excludedVG;
        total.degenerate += pst[s].degenerate;
    }
    if (stats) *stats = total;
    return merged;
}

}  // namespace fc
