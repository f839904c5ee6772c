// Wall-model cell list (see fc_wall.hpp).
//
// Two passes over the (y, z) rows of the grid, both dealt to host threads
// through one atomic row counter (rows are independent, so any thread may
// take any row):
//   1. count the candidates of every row: fluid cells with at least one
//      solid pull neighbour;
//   2. after an exclusive prefix sum of the counts, evaluate every candidate
//      into its slot of one preallocated entry array (row offset + rank in
//      the row) with a status: listed, vane-adjacent, degenerate normal.
// A final stable pass copies the listed entries into the SoA output. The
// slots follow the x-fastest cell order, so the list is the sequential
// scan's list whatever the thread count.
#include "fc_wall.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <thread>

namespace fc {
namespace {

constexpr float kDegenerateEps = 1e-3f;  // wall_model.cpp:18

enum : std::uint8_t { kListed = 0, kNearVane = 1, kDegenerate = 2 };

struct Entry {
    long long cell, sample;
    float n[3], dist;
    std::uint32_t links;
    std::uint8_t status;
};

class Lattice {
public:
    Lattice(const Dims& d, const std::vector<std::uint8_t>& active, const std::vector<std::uint8_t>& clean)
        : d_(d), active_(active), clean_(clean) {}

    long long index(int x, int y, int z) const {
        return static_cast<long long>(x) +
               static_cast<long long>(d_.nx) * (static_cast<long long>(y) + static_cast<long long>(d_.ny) * z);
    }
    int wrapZ(int z) const { return ((z % d_.nz) + d_.nz) % d_.nz; }
    bool inPlane(int x, int y) const { return x >= 0 && x < d_.nx && y >= 0 && y < d_.ny; }
    // Off-grid x / y is not solid; z is periodic.
    bool solidIn(const std::vector<std::uint8_t>& f, int x, int y, int z) const {
        return inPlane(x, y) && f[index(x, y, wrapZ(z))] == kSolid;
    }
    bool fluid(long long c) const { return active_[c] == kFluid; }

    // Bit q set when the pull source x - c_q is solid in the live flags.
    std::uint32_t solidLinks(int x, int y, int z) const {
        std::uint32_t m = 0;
        for (int q = 1; q < kQ; ++q)
            if (solidIn(active_, x - kCx[q], y - kCy[q], z - kCz[q])) m |= 1u << q;
        return m;
    }

    // The per-cell rule of wall_model.cpp:55-165, bit for bit (the lists are
    // compared entry by entry with the reference's).
    Entry evaluate(int x, int y, int z, std::uint32_t links, int sampleCells) const {
        Entry e{};
        e.cell = index(x, y, z);
        e.links = links;
        // A link whose source is solid only because of a vane is not a foil
        // wall: such cells stay off the list.
        for (int q = 1; q < kQ; ++q)
            if ((links >> q & 1u) && !solidIn(clean_, x - kCx[q], y - kCy[q], z - kCz[q])) {
                e.status = kNearVane;
                return e;
            }
        // Outward normal: minus the first moment of the clean solid cells in
        // the (2R+1)^3 box (integer sums, exact in FP32).
        float m[3] = {0.0f, 0.0f, 0.0f};
        const int r = kWallNormalRadius;
        for (int dz = -r; dz <= r; ++dz)
            for (int dy = -r; dy <= r; ++dy)
                for (int dx = -r; dx <= r; ++dx)
                    if (solidIn(clean_, x + dx, y + dy, z + dz)) {
                        m[0] += static_cast<float>(dx);
                        m[1] += static_cast<float>(dy);
                        m[2] += static_cast<float>(dz);
                    }
        const float len = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
        if (len < kDegenerateEps) {
            e.status = kDegenerate;
            return e;
        }
        for (int a = 0; a < 3; ++a) e.n[a] = -m[a] / len;
        // Velocity sample: the first fluid cell along the normal, tried from
        // sampleCells cells out down to 1; the wall cell itself otherwise.
        int off[3] = {0, 0, 0};
        e.sample = e.cell;
        for (int k = sampleCells; k >= 1; --k) {
            const int o[3] = {static_cast<int>(std::lround(e.n[0] * k)), static_cast<int>(std::lround(e.n[1] * k)),
                              static_cast<int>(std::lround(e.n[2] * k))};
            if (!inPlane(x + o[0], y + o[1])) continue;
            const long long c = index(x + o[0], y + o[1], wrapZ(z + o[2]));
            if (!fluid(c)) continue;
            e.sample = c;
            off[0] = o[0];
            off[1] = o[1];
            // spanwise offset as its minimal periodic image
            int oz = wrapZ(z + o[2]) - z;
            if (oz > d_.nz / 2) oz -= d_.nz;
            if (oz < -d_.nz / 2) oz += d_.nz;
            off[2] = oz;
            break;
        }
        const float along = static_cast<float>(off[0]) * e.n[0] + static_cast<float>(off[1]) * e.n[1] +
                            static_cast<float>(off[2]) * e.n[2];
        e.dist = 0.5f + std::max(along, 0.0f);
        e.status = kListed;
        return e;
    }

    const Dims& dims() const { return d_; }

private:
    const Dims& d_;
    const std::vector<std::uint8_t>& active_;
    const std::vector<std::uint8_t>& clean_;
};

// Runs body(row) for rows [0, rows) on up to `threads` host threads that
// pull rows from one shared counter.
template <class F>
void forEachRow(long long rows, int threads, const F& body) {
    std::atomic<long long> next{0};
    auto worker = [&]() {
        for (long long r = next.fetch_add(1); r < rows; r = next.fetch_add(1)) body(r);
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) pool.emplace_back(worker);
    worker();
    for (auto& t : pool) t.join();
}

}  // namespace

WallList buildWallCellList(const Dims& d, const std::vector<std::uint8_t>& active,
                           const std::vector<std::uint8_t>& clean, int sampleCells,
                           WallListStats* stats) {
    WallList out;
    WallListStats total;
    const std::size_t n = static_cast<std::size_t>(d.cells());
    if (n == 0 || active.size() != n || clean.size() != n) {
        if (stats) *stats = total;
        return out;
    }
    const Lattice lat(d, active, clean);
    const long long rows = static_cast<long long>(d.ny) * d.nz;
    const int threads = static_cast<int>(
        std::min<long long>(rows, std::max(1u, std::thread::hardware_concurrency())));

    // pass 1: candidates per row
    std::vector<long long> offset(static_cast<std::size_t>(rows) + 1, 0);
    forEachRow(rows, threads, [&](long long r) {
        const int y = static_cast<int>(r % d.ny), z = static_cast<int>(r / d.ny);
        long long c = 0;
        for (int x = 0; x < d.nx; ++x)
            if (lat.fluid(lat.index(x, y, z)) && lat.solidLinks(x, y, z)) ++c;
        offset[static_cast<std::size_t>(r) + 1] = c;
    });
    for (long long r = 0; r < rows; ++r) offset[r + 1] += offset[r];

    // pass 2: evaluate every candidate into its slot
    std::vector<Entry> slot(static_cast<std::size_t>(offset[rows]));
    forEachRow(rows, threads, [&](long long r) {
        const int y = static_cast<int>(r % d.ny), z = static_cast<int>(r / d.ny);
        long long k = offset[r];
        for (int x = 0; x < d.nx; ++x) {
            if (!lat.fluid(lat.index(x, y, z))) continue;
            const std::uint32_t links = lat.solidLinks(x, y, z);
            if (links) slot[static_cast<std::size_t>(k++)] = lat.evaluate(x, y, z, links, sampleCells);
        }
    });

    // stable compaction of the listed entries
    for (const Entry& e : slot) {
        if (e.status == kNearVane) {
            ++total.excludedVG;
            continue;
        }
        if (e.status == kDegenerate) {
            ++total.degenerate;
            continue;
        }
        out.cellIdx.push_back(e.cell);
        out.sampleIdx.push_back(e.sample);
        out.nx.push_back(e.n[0]);
        out.ny.push_back(e.n[1]);
        out.nz.push_back(e.n[2]);
        out.sampleDist.push_back(e.dist);
        out.linkMask.push_back(e.links);
        ++total.listed;
    }
    if (stats) *stats = total;
    return out;
}

}  // namespace fc
