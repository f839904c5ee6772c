// Host + device shared constants: D3Q19 tables, cell flags, the per-cell
// kind byte and the HBM data layout of the in-place (AA) population store.
// Plain C++ (no HIP types) so host translation units compiled with g++ and
// device translation units compiled with hipcc see identical definitions.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define FC_HD __host__ __device__
#else
#define FC_HD
#endif

namespace fc {

inline constexpr int kQ = 19;

// D3Q19 velocities, opposites and the y / z mirrors used by the free-slip
// walls; frozen ordering of the reference (src/sim/lbm_core.cuh:34-61).
inline constexpr int kCx[kQ] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0};
inline constexpr int kCy[kQ] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1};
inline constexpr int kCz[kQ] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1};
inline constexpr int kOpp[kQ] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
inline constexpr int kMirY[kQ] = {0, 1, 2, 4, 3, 5, 6, 13, 14, 9, 10, 18, 17, 7, 8, 15, 16, 12, 11};
inline constexpr int kMirZ[kQ] = {0, 1, 2, 3, 4, 6, 5, 7, 8, 15, 16, 17, 18, 13, 14, 9, 10, 11, 12};

// Cell flags, numeric values frozen with the reference (lbm_core.cuh:114-130).
enum : std::uint8_t {
    kFluid = 0,
    kSolid = 1,
    kInlet = 2,
    kOutlet = 3,
    kSlipTop = 4,
    kSlipBottom = 5,
    kSlipFront = 6,
    kSlipBack = 7,
    kInterface = 8,
};

// The device kind byte per cell: low nibble = the cell flag above, high bits
// are derived at flag upload so the step kernel never reads neighbour flags
// for a bulk fluid cell:
//   kKindSpecial : fluid cell with at least one solid or slip pull neighbour
//                  (per-link source select needed);
//   kKindFeed    : the +x neighbour is an Outlet cell (this cell publishes its
//                  post-collision populations for the zero-gradient copy);
//   kKindWall    : fluid cell carrying a wall-model slip velocity.
inline constexpr std::uint8_t kKindFlagMask = 0x0F;
inline constexpr std::uint8_t kKindSpecial = 0x10;
inline constexpr std::uint8_t kKindFeed = 0x20;
inline constexpr std::uint8_t kKindWall = 0x40;

// Population storage formats (the C-ABI's FC_STORAGE_* values).
enum Storage : int {
    kStoreF32 = 0,
    kStoreF16 = 1,
    kStoreBF16 = 2,
    kStoreF64 = 3,
    kStoreF16C = 4,
};

inline constexpr int storageBytes(int s) {
    return s == kStoreF32 ? 4 : (s == kStoreF64 ? 8 : 2);
}

// In-place AA streaming: one population array, f[q * ncells + cell],
// cell = x + nx*(y + ny*z). The post-collision population f*_q(c) of cell c
// lives, depending on the layout state of the array, at
//   layout L ("local")   : slot (opp(q), c)
//   layout S ("scatter") : slot (q, c + e_q)
// A step reads its inputs in the current layout and writes its outputs in
// the other one: from L it reads (opp q, c - e_q) / writes (q, c + e_q);
// from S it reads (q, c) / writes (opp q, c). Each slot is read and written
// by the same cell within a step, so the update is race free in place.
enum Layout : int { kLayoutL = 0, kLayoutS = 1 };

struct Dims {
    int nx = 0, ny = 0, nz = 0;
    FC_HD long long cells() const { return (long long)nx * ny * nz; }
};

// Wave-row decomposition of the step kernels: one 64-lane wavefront owns
// 64 consecutive x cells of one (y, z) row; a workgroup holds 4 waves.
inline constexpr int kWave = 64;
inline constexpr int kWavesPerBlock = 4;
inline constexpr int kBlock = kWave * kWavesPerBlock;
inline constexpr int kXcds = 8;  // MI355X: 8 XCDs, workgroups dealt round-robin

// Integer hash of the stochastic rounding's random bits (lowbias32:
// xorshift-multiply rounds with the constants of C. Wellons' hash-prospector
// search) of a value and a seed; constexpr, so host and device share it.
constexpr unsigned roundHash(unsigned c, unsigned seed) {
    unsigned x = c ^ seed;
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    return x;
}

}  // namespace fc
