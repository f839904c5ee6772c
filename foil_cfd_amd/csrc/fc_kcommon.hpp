// Device helpers shared by the FoilCFD kernel translation units: the kernel
// argument block of one lattice level, the wave-row decomposition with its
// scalar row offsets, layout-aware reads of post-collision populations and
// deterministic FP64 wave / block reductions.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

#include "fc_device.hpp"
#include "fc_kernels.hpp"

namespace fc {
namespace {

// Kernel arguments of one lattice level (POD, passed by value).
struct KArgs {
    void* f;
    const std::uint8_t* kind;
    void* feed;
    float* rho;
    float* u;
    float* v;
    float* w;
    const std::uint16_t* uwx;
    const std::uint16_t* uwy;
    const std::uint16_t* uwz;
    long long ncells;
    int nx, ny, nz;
    int nseg;     // 64-cell segments per row
    int nwaves;   // nseg * ny * nz
    int nblocks;  // ceil(nwaves / kWavesPerBlock)
    int chunk;    // blocks per XCD after the remap
    // Reciprocal multipliers for the task -> (row, z) divisions: n / d as
    // (n * m) >> 40 with m = ceil(2^40 / d), exact for n < 2^24, d < 2^16
    // (0 when the grid exceeds those ranges: plain division).
    unsigned long long segMagic, nyMagic;
    float tau0, lam, pre, uIn;
    int writeMacro;
    // 1: population element indices need 64 bits (kQ * ncells >= 2^32); the
    // plain-wall step then takes the 64-bit slot select. Also set on request
    // (the wide step variant) so that path runs on small grids.
    int wideIndex;
    // Stochastic rounding of the 16-bit population stores (1) or round to
    // nearest (0), and this launch's seed for the random bits.
    int stochastic;
    unsigned roundSeed;
};

// Random bits for the stochastic rounding of one cell's stores in one step:
// roundHash (fc_layout.hpp) of the cell index and the launch seed, then a
// Weyl sequence over the 19 directions (the storage codecs use the top
// bits). A pure function of (cell, direction, seed), so every kernel variant
// that stores a cell's populations draws the same bits.
__device__ inline unsigned roundBits(unsigned h, int q) { return h + static_cast<unsigned>(q) * 0x9E3779B9u; }

inline unsigned long long divMagic(long long d) {
    return d > 0 && d < (1 << 16) ? ((1ull << 40) + static_cast<unsigned long long>(d) - 1) /
                                        static_cast<unsigned long long>(d)
                                  : 0ull;
}

inline void setDivMagic(KArgs& a) {
    const bool ok = a.nwaves < (1 << 24);
    a.segMagic = ok ? divMagic(a.nseg) : 0ull;
    a.nyMagic = ok ? divMagic(a.ny) : 0ull;
    if (!a.segMagic || !a.nyMagic) a.segMagic = a.nyMagic = 0ull;
}

// n / d for a wave-uniform task index (reciprocal multiply when available).
__device__ inline int taskDiv(int n, int d, unsigned long long m) {
    return m ? static_cast<int>((static_cast<unsigned long long>(n) * m) >> 40) : n / d;
}

KArgs makeArgs(const LatticeView& lat) {
    KArgs a{};
    a.f = lat.f;
    a.kind = lat.kind;
    a.feed = lat.feed;
    a.rho = lat.rho;
    a.u = lat.u;
    a.v = lat.v;
    a.w = lat.w;
    a.uwx = lat.uwx;
    a.uwy = lat.uwy;
    a.uwz = lat.uwz;
    a.ncells = lat.dims.cells();
    a.nx = lat.dims.nx;
    a.ny = lat.dims.ny;
    a.nz = lat.dims.nz;
    a.nseg = (a.nx + kWave - 1) / kWave;
    a.nwaves = a.nseg * a.ny * a.nz;
    a.nblocks = (a.nwaves + kWavesPerBlock - 1) / kWavesPerBlock;
    a.chunk = (a.nblocks + kXcds - 1) / kXcds;
    a.wideIndex = a.ncells * kQ >= (1ll << 32) ? 1 : 0;
    setDivMagic(a);
    return a;
}

// Position of this lane in the wave-row decomposition plus the 3 x 3 table
// of row start indices for y-1..y+1, z-1..z+1 (z periodic; y and x periodic
// only when requested). Row starts of rows outside a non-periodic y range
// are flagged by yOk so boundary cells can skip out-of-domain scatter writes.
struct RowPos {
    int x, y, z;
    int xm, xp;
    int R[3][3];
    bool yOk[3];
};

// Lane position of row-segment task gw (wave-uniform; segment of 64 x of
// row gw / nseg). False for lanes past the end of the row.
template <bool PX, bool PY>
__device__ inline bool taskPos(const KArgs& a, int gw, RowPos& p) {
    const int row = taskDiv(gw, a.nseg, a.segMagic);
    const int seg = gw - row * a.nseg;
    const int z = taskDiv(row, a.ny, a.nyMagic);
    const int y = row - z * a.ny;
    const int x = seg * kWave + static_cast<int>(threadIdx.x & (kWave - 1));
    p.x = x;
    p.y = y;
    p.z = z;
    p.xm = x - 1;
    p.xp = x + 1;
    if (PX) {
        if (p.xm < 0) p.xm += a.nx;
        if (p.xp >= a.nx) p.xp -= a.nx;
    }
    int ys[3] = {y - 1, y, y + 1};
    p.yOk[0] = p.yOk[1] = p.yOk[2] = true;
    if (PY) {
        if (ys[0] < 0) ys[0] += a.ny;
        if (ys[2] >= a.ny) ys[2] -= a.ny;
    } else {
        p.yOk[0] = ys[0] >= 0;
        p.yOk[2] = ys[2] < a.ny;
        if (!p.yOk[0]) ys[0] = 0;
        if (!p.yOk[2]) ys[2] = a.ny - 1;
    }
    const int zs[3] = {z == 0 ? a.nz - 1 : z - 1, z, z == a.nz - 1 ? 0 : z + 1};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) p.R[i][j] = (ys[i] + a.ny * zs[j]) * a.nx;
    return x < a.nx;
}

template <bool PX, bool PY>
__device__ inline bool rowPos(const KArgs& a, RowPos& p) {
    const int hb = blockIdx.x;
    const int lb = (hb & (kXcds - 1)) * a.chunk + (hb >> 3);
    if (lb >= a.nblocks) return false;
    const int wv = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) >> 6);
    const int gw = lb * kWavesPerBlock + wv;
    if (gw >= a.nwaves) return false;
    return taskPos<PX, PY>(a, gw, p);
}

// Compile-time loop: f(std::integral_constant<int, i>) for i = 0 .. N-1, so
// direction tables index as constants inside the body.
template <class F, int... Is>
__device__ inline void staticForImpl(F&& f, std::integer_sequence<int, Is...>) {
    (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, class F>
__device__ inline void staticFor(F&& f) {
    staticForImpl(f, std::make_integer_sequence<int, N>{});
}

// Cell index of (x + ex, y + ey, z + ez) for compile-time offsets.
__device__ inline unsigned nb(const RowPos& p, int ex, int ey, int ez) {
    const int xx = ex < 0 ? p.xm : (ex > 0 ? p.xp : p.x);
    return static_cast<unsigned>(xx + p.R[ey + 1][ez + 1]);
}

// True when (x + ex, y + ey) is inside a non-periodic domain (z wraps).
template <bool PX>
__device__ inline bool inDomain(const KArgs& a, const RowPos& p, int ex, int ey) {
    bool ok = p.yOk[ey + 1];
    if (!PX) {
        const int xx = p.x + ex;
        ok = ok && xx >= 0 && xx < a.nx;
    }
    return ok;
}

// ---------------------------------------------------------------------------
// Population accessors of the step kernel: slot (q, cell (x+ex, y+ey, z+ez))
// for compile-time offsets.
//
// PopPtr: flat 64-bit addressing, any array size.
// PopBuf: raw buffer instructions over one wave-uniform descriptor. The
//   direction slab and the row start are a scalar byte offset (SGPR soffset),
//   only the x byte offset lives in a vector register, so an access costs one
//   scalar add and no per-lane address arithmetic. Every offset must stay
//   below 2^31 bytes, so the host only selects it for arrays under 2 GiB.
// ---------------------------------------------------------------------------
// Storage bits of a value: the 16-bit formats round stochastically with the
// random bits r when sto is set (else, and for 32 / 64 bits, to nearest).
template <class S>
__device__ inline typename S::Bits roundedBits(typename S::C v, bool sto, unsigned r) {
    if constexpr (sizeof(typename S::T) == 2)
        return S::toBitsR(v, sto, r);
    else
        return S::toBits(v);
}

template <class S>
struct PopPtr {
    using T = typename S::T;
    using C = typename S::C;
    T* f;
    long long n;
    const RowPos* p;
    __device__ PopPtr(const KArgs& a, const RowPos& rp) : f(static_cast<T*>(a.f)), n(a.ncells), p(&rp) {}
    __device__ C ld(int q, int ex, int ey, int ez) const {
        return S::load(f + static_cast<long long>(q) * n + nb(*p, ex, ey, ez));
    }
    __device__ void st(int q, int ex, int ey, int ez, C v) const {
        S::store(f + static_cast<long long>(q) * n + nb(*p, ex, ey, ez), v);
    }
    __device__ void stBits(int q, int ex, int ey, int ez, typename S::Bits b) const {
        reinterpret_cast<typename S::Bits*>(f)[static_cast<long long>(q) * n + nb(*p, ex, ey, ez)] = b;
    }
};

// Cache-policy bits of the population buffer accesses (gfx950 CPol: bit 0
// sc0, bit 1 nt, bit 4 sc1); 0 = the default policy.
#ifndef FC_POP_LOAD_POLICY
#define FC_POP_LOAD_POLICY 0
#endif
#ifndef FC_POP_STORE_POLICY
#define FC_POP_STORE_POLICY 0
#endif
inline constexpr int kPopLoadPolicy = FC_POP_LOAD_POLICY;
inline constexpr int kPopStorePolicy = FC_POP_STORE_POLICY;

template <class S>
struct PopBuf {
    using T = typename S::T;
    using C = typename S::C;
    using Bits = typename S::Bits;
    static constexpr unsigned kSz = sizeof(T);
    __amdgpu_buffer_rsrc_t rs;
    unsigned slab;       // bytes per direction slab (ncells * sizeof(T))
    unsigned vx[3];      // per-lane byte offsets of x-1, x, x+1
    unsigned row[3][3];  // wave-uniform byte offsets of the 9 row starts
    PopBuf() = default;
    __device__ PopBuf(const KArgs& a, const RowPos& p) {
        // 0x00020000: raw (untyped) buffer descriptor word 3 of the gfx9 family
        rs = __builtin_amdgcn_make_buffer_rsrc(a.f, 0, 0x7FFFFFFF, 0x00020000);
        slab = static_cast<unsigned>(a.ncells) * kSz;
        vx[0] = static_cast<unsigned>(p.xm) * kSz;
        vx[1] = static_cast<unsigned>(p.x) * kSz;
        vx[2] = static_cast<unsigned>(p.xp) * kSz;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) row[i][j] = static_cast<unsigned>(p.R[i][j]) * kSz;
    }
    __device__ C ld(int q, int ex, int ey, int ez) const {
        const int vo = static_cast<int>(vx[ex + 1]);
        const int so = static_cast<int>(static_cast<unsigned>(q) * slab + row[ey + 1][ez + 1]);
        if constexpr (sizeof(Bits) == 2)
            return S::fromBits(static_cast<Bits>(__builtin_amdgcn_raw_buffer_load_b16(rs, vo, so, kPopLoadPolicy)));
        else
            return S::fromBits(static_cast<Bits>(__builtin_amdgcn_raw_buffer_load_b32(rs, vo, so, kPopLoadPolicy)));
    }
    // Raw storage bits of a slot (decoded later with S::fromBits), so a load
    // can stay in flight across other work.
    __device__ Bits ldBits(int q, int ex, int ey, int ez) const {
        const int vo = static_cast<int>(vx[ex + 1]);
        const int so = static_cast<int>(static_cast<unsigned>(q) * slab + row[ey + 1][ez + 1]);
        if constexpr (sizeof(Bits) == 2)
            return static_cast<Bits>(__builtin_amdgcn_raw_buffer_load_b16(rs, vo, so, kPopLoadPolicy));
        else
            return static_cast<Bits>(__builtin_amdgcn_raw_buffer_load_b32(rs, vo, so, kPopLoadPolicy));
    }
    __device__ void st(int q, int ex, int ey, int ez, C v) const { stBits(q, ex, ey, ez, S::toBits(v)); }
    __device__ void stBits(int q, int ex, int ey, int ez, Bits b) const {
        const int vo = static_cast<int>(vx[ex + 1]);
        const int so = static_cast<int>(static_cast<unsigned>(q) * slab + row[ey + 1][ez + 1]);
        if constexpr (sizeof(Bits) == 2)
            __builtin_amdgcn_raw_buffer_store_b16(b, rs, vo, so, kPopStorePolicy);
        else
            __builtin_amdgcn_raw_buffer_store_b32(b, rs, vo, so, kPopStorePolicy);
    }

};

// Post-collision population f*_q of cell (x,y,z) in the given layout, as a
// deviation H (unscaled by K). Used by the readout kernels.
template <class S>
__device__ inline typename S::C postH(const KArgs& a, int layout, int q, int x, int y, int z) {
    using T = typename S::T;
    const T* const f = static_cast<const T*>(a.f);
    if (layout == kLayoutL) {
        const long long c = x + static_cast<long long>(a.nx) * (y + static_cast<long long>(a.ny) * z);
        return S::load(f + static_cast<long long>(kOpp[q]) * a.ncells + c);
    }
    int xx = x + kCx[q], yy = y + kCy[q], zz = z + kCz[q];
    if (xx < 0) xx += a.nx;
    if (xx >= a.nx) xx -= a.nx;
    if (yy < 0) yy += a.ny;
    if (yy >= a.ny) yy -= a.ny;
    if (zz < 0) zz += a.nz;
    if (zz >= a.nz) zz -= a.nz;
    const long long c = xx + static_cast<long long>(a.nx) * (yy + static_cast<long long>(a.ny) * zz);
    return S::load(f + static_cast<long long>(q) * a.ncells + c);
}

// Wave64 FP64 sum via cross-lane shuffles.
__device__ inline double waveSum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
    return v;
}

// Wave64 reductions of FP32 / int partials (the full sum lands in lane 0).
__device__ inline float waveSumF(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
    return v;
}
__device__ inline int waveSumI(int v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
    return v;
}
__device__ inline int waveMaxI(int v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_down(v, off, kWave));
    return v;
}

// Deterministic block reduction of NV doubles per thread (block = kBlock).
template <int NV>
__device__ inline void blockSum(double (&v)[NV], double* out) {
    __shared__ double sm[NV][kWavesPerBlock];
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        const double s = waveSum(v[i]);
        if (lane == 0) sm[i][wv] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            double t = 0;
            for (int j = 0; j < kWavesPerBlock; ++j) t += sm[i][j];
            out[i] = t;
        }
    }
}

}  // namespace
}  // namespace fc
