// Tracer particles for gfx950: lowbias32 seeding and RK2 (midpoint)
// advection through the trilinearly sampled macroscopic velocity field, with
// outflow / out-of-domain / solid-entry / lifetime respawn at the inlet plane
// or in the fluid volume. Same rules, constants and operation order as the
// reference (src/render/particles.cu:48-376); floating-point contraction is
// switched off so every particle is bit-identical to the CPU restatement.
#include <hip/hip_runtime.h>

#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>

#include "fc_kcommon.hpp"
#include "fc_refine_math.hpp"

#pragma clang fp contract(off)

namespace fc {
namespace {

__device__ inline unsigned hashU32(unsigned x) {  // lowbias32
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__device__ inline float rand01(unsigned& s) {
    s = hashU32(s);
    return static_cast<float>(s & 0x00FFFFFFu) * (1.0f / 16777216.0f);
}

__device__ inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ inline long long cellAt(const Dims& d, int x, int y, int z) {
    return static_cast<long long>(x) +
           static_cast<long long>(d.nx) * (static_cast<long long>(y) + static_cast<long long>(d.ny) * z);
}

struct V3 {
    float x, y, z;
};

__device__ inline V3 lerp3(V3 a, V3 b, float t) {
    return V3{a.x + (b.x - a.x) * t, a.y + (b.y - a.y) * t, a.z + (b.z - a.z) * t};
}

// Velocity sources of the sampler: the three macro arrays, or one packed
// (u, v, w, 0) vector per cell so that a corner is a single 16-byte read
// instead of three scattered 4-byte reads. Same values either way.
struct VelSoA {
    const float* u;
    const float* v;
    const float* w;
    const std::uint8_t* kind;
    __device__ V3 at(long long c) const { return V3{u[c], v[c], w[c]}; }
    __device__ bool solid(long long c) const { return kind && (kind[c] & kKindFlagMask) == kSolid; }
};
// The packed vector's fourth component marks solid cells (1), so the final
// solid-entry test reads a vector the two samples most often just fetched.
struct VelPacked {
    const float4* vel;
    __device__ V3 at(long long c) const {
        const float4 t = vel[c];
        return V3{t.x, t.y, t.z};
    }
    __device__ bool solid(long long c) const { return vel[c].w != 0.0f; }
};

template <class VS>
__device__ inline V3 velAt(const VS& vs, const Dims& d, int x, int y, int z) {
    int zz = z;
    if (zz < 0) zz += d.nz;
    else if (zz >= d.nz) zz -= d.nz;
    return vs.at(cellAt(d, clampi(x, 0, d.nx - 1), clampi(y, 0, d.ny - 1), zz));
}

// Trilinear sample at p - 0.5 (cell centres at +0.5), x / y clamped, z periodic.
template <class VS>
__device__ inline V3 sampleVel(const VS& vs, const Dims& d, V3 p) {
    const float gx = fminf(fmaxf(p.x - 0.5f, 0.0f), static_cast<float>(d.nx) - 1.0001f);
    const float gy = fminf(fmaxf(p.y - 0.5f, 0.0f), static_cast<float>(d.ny) - 1.0001f);
    float gz = p.z - 0.5f;
    if (gz < 0.0f) gz += static_cast<float>(d.nz);
    const int x0 = static_cast<int>(gx), y0 = static_cast<int>(gy), z0 = static_cast<int>(gz);
    const float fx = gx - static_cast<float>(x0);
    const float fy = gy - static_cast<float>(y0);
    const float fz = gz - static_cast<float>(z0);
    const V3 c000 = velAt(vs, d, x0, y0, z0), c100 = velAt(vs, d, x0 + 1, y0, z0);
    const V3 c010 = velAt(vs, d, x0, y0 + 1, z0), c110 = velAt(vs, d, x0 + 1, y0 + 1, z0);
    const V3 c001 = velAt(vs, d, x0, y0, z0 + 1), c101 = velAt(vs, d, x0 + 1, y0, z0 + 1);
    const V3 c011 = velAt(vs, d, x0, y0 + 1, z0 + 1), c111 = velAt(vs, d, x0 + 1, y0 + 1, z0 + 1);
    const V3 x00 = lerp3(c000, c100, fx), x10 = lerp3(c010, c110, fx);
    const V3 x01 = lerp3(c001, c101, fx), x11 = lerp3(c011, c111, fx);
    return lerp3(lerp3(x00, x10, fy), lerp3(x01, x11, fy), fz);
}

__device__ inline bool finite3(V3 a) {
    return __builtin_isfinite(a.x) && __builtin_isfinite(a.y) && __builtin_isfinite(a.z);
}

__device__ inline V3 inletSpawn(const Dims& d, unsigned& r) {
    const float a = 0.5f + 1.5f * rand01(r);
    const float b = static_cast<float>(d.ny) * rand01(r);
    const float c = static_cast<float>(d.nz) * rand01(r);
    return V3{a, b, c};
}

__device__ inline bool isSolid(const std::uint8_t* kind, long long c) {
    return (kind[c] & kKindFlagMask) == kSolid;
}

__device__ inline V3 volumeSpawn(const Dims& d, const std::uint8_t* kind, unsigned& r) {
    for (int attempt = 0; attempt < 4; ++attempt) {
        const float a = static_cast<float>(d.nx) * rand01(r);
        const float b = static_cast<float>(d.ny) * rand01(r);
        const float c = static_cast<float>(d.nz) * rand01(r);
        if (!kind) return V3{a, b, c};
        const long long cell = cellAt(d, clampi(static_cast<int>(a), 0, d.nx - 1),
                                      clampi(static_cast<int>(b), 0, d.ny - 1),
                                      clampi(static_cast<int>(c), 0, d.nz - 1));
        if (!isSolid(kind, cell)) return V3{a, b, c};
    }
    return inletSpawn(d, r);
}

__global__ void __launch_bounds__(256) seedKernel(float4* pos, int count, Dims d, unsigned seed) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    unsigned r = hashU32(seed ^ (static_cast<unsigned>(i) * 0x9E3779B9u));
    const float a = static_cast<float>(d.nx) * rand01(r);
    const float b = static_cast<float>(d.ny) * rand01(r);
    const float c = static_cast<float>(d.nz) * rand01(r);
    const float e = rand01(r);
    pos[i] = make_float4(a, b, c, e);
}

// (u, v, w, solid mark) of every cell into one 16-byte vector.
__global__ void __launch_bounds__(256) packVelocityKernel(const float* u, const float* v, const float* w,
                                                          const std::uint8_t* kind, float4* vel, long long n) {
    const long long c = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c < n) vel[c] = make_float4(u[c], v[c], w[c], kind && isSolid(kind, c) ? 1.0f : 0.0f);
}

// Sort key of a particle: the Morton (Z-order) index of the cell it is in
// (clamped into the grid; non-finite positions sort to cell 0), so particles
// that share trilinear corners in x, y or z sort close together and a run of
// the sorted list covers a compact block of cells. Grids up to 1024 cells
// per axis (30-bit keys); larger grids fall back to the plain cell index.
__host__ __device__ inline unsigned spread3(unsigned v) {  // bit i -> bit 3i, 10 bits
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__host__ __device__ inline bool mortonKeys(const Dims& d) { return d.nx <= 1024 && d.ny <= 1024 && d.nz <= 1024; }

__global__ void __launch_bounds__(256) particleKeyKernel(const float4* pos, int count, Dims d, unsigned* key,
                                                         int* slot) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float4 p = pos[i];
    const bool ok = __builtin_isfinite(p.x) && __builtin_isfinite(p.y) && __builtin_isfinite(p.z);
    const int x = ok ? clampi(static_cast<int>(p.x), 0, d.nx - 1) : 0;
    const int y = ok ? clampi(static_cast<int>(p.y), 0, d.ny - 1) : 0;
    const int z = ok ? clampi(static_cast<int>(p.z), 0, d.nz - 1) : 0;
    key[i] = mortonKeys(d) ? spread3(static_cast<unsigned>(x)) | (spread3(static_cast<unsigned>(y)) << 1) |
                                 (spread3(static_cast<unsigned>(z)) << 2)
                           : static_cast<unsigned>(cellAt(d, x, y, z));
    slot[i] = i;
}

// One particle per lane. With `order` (slots sorted by cell), lane t of the
// XCD-contiguous block range advances slot order[t]: neighbouring lanes and
// the blocks of one XCD then sample neighbouring cells, so the trilinear
// gathers hit lines already in L2 instead of a new line per corner. The
// update of a slot (and its random stream, seeded by the slot) does not
// depend on the order, so both orders give identical particles.
template <class VS>
__global__ void __launch_bounds__(256) advectKernel(float4* pos, int count, VS vs, const std::uint8_t* kind, Dims d,
                                                    ParticleParams prm, const int* order) {
    int i;
    if (order) {
        const long long blocks = (static_cast<long long>(count) + 255) / 256;
        const long long t = xcdBlock(blocks) * 256 + threadIdx.x;
        if (t >= count) return;
        i = order[t];
    } else {
        i = blockIdx.x * blockDim.x + threadIdx.x;
        if (i >= count) return;
    }
    unsigned r = hashU32(prm.seed ^ (static_cast<unsigned>(i) * 0x9E3779B9u));
    const float4 pk = pos[i];
    V3 p{pk.x, pk.y, pk.z};
    float age = pk.w;
    const float nzf = static_cast<float>(d.nz);
    bool blew = !finite3(p);
    if (!blew && prm.dtSteps > 0.0f) {
        const V3 u1 = sampleVel(vs, d, p);
        V3 pm{p.x + 0.5f * prm.dtSteps * u1.x, p.y + 0.5f * prm.dtSteps * u1.y, p.z + 0.5f * prm.dtSteps * u1.z};
        if (pm.z < 0.0f) pm.z += nzf;
        else if (pm.z >= nzf) pm.z -= nzf;
        const V3 u2 = sampleVel(vs, d, pm);
        p = V3{p.x + prm.dtSteps * u2.x, p.y + prm.dtSteps * u2.y, p.z + prm.dtSteps * u2.z};
        blew = !finite3(p) || !finite3(u2);
        age -= prm.ageRate * prm.dtSteps;
    }
    if (!blew) {
        if (p.z < 0.0f) p.z += nzf;
        else if (p.z >= nzf) p.z -= nzf;
    }
    const bool outflow = blew || p.x >= static_cast<float>(d.nx) - 1.0f;
    const bool oob = !outflow && (p.x < 0.0f || p.y < 0.0f || p.y >= static_cast<float>(d.ny));
    const int cx = clampi(static_cast<int>(p.x), 0, d.nx - 1);
    const int cy = clampi(static_cast<int>(p.y), 0, d.ny - 1);
    const int cz = clampi(static_cast<int>(p.z), 0, d.nz - 1);
    const bool solid = !outflow && !oob && vs.solid(cellAt(d, cx, cy, cz));
    if (outflow || oob || solid || age <= 0.0f) {
        const bool toInlet = outflow || oob || rand01(r) < prm.inletFraction;
        p = toInlet ? inletSpawn(d, r) : volumeSpawn(d, kind, r);
        age = 0.6f + 0.4f * rand01(r);
    }
    pos[i] = make_float4(p.x, p.y, p.z, age);
}

}  // namespace

int launchParticleSeed(float* pos4, int count, const Dims& d, unsigned seed, void* stream) {
    if (!pos4 || count <= 0 || d.cells() <= 0) return static_cast<int>(hipErrorInvalidValue);
    float4* pos = reinterpret_cast<float4*>(pos4);
    seedKernel<<<(count + 255) / 256, 256, 0, static_cast<hipStream_t>(stream)>>>(pos, count, d, seed);
    return static_cast<int>(hipGetLastError());
}

namespace {

int keyBits(const Dims& d) {
    if (mortonKeys(d)) {  // 3 x the bits of the longest axis
        int b = 1;
        const int m = d.nx > d.ny ? (d.nx > d.nz ? d.nx : d.nz) : (d.ny > d.nz ? d.ny : d.nz);
        while ((1 << b) < m) ++b;
        return 3 * b;
    }
    int b = 1;
    while (b < 32 && (1ll << b) < d.cells()) ++b;
    return b;
}

// Layout of the order scratch: keys, alternate keys, slots, alternate slots
// (count each, 256-byte aligned), then rocPRIM's radix-sort storage.
struct OrderScratch {
    unsigned *key, *keyAlt;
    int *slot, *slotAlt;
    void* sortTmp;
    std::size_t sortBytes;
};

std::size_t alignUp(std::size_t b) { return (b + 255) & ~static_cast<std::size_t>(255); }

std::size_t sortTempBytes(int count, int bits) {
    std::size_t bytes = 0;
    rocprim::double_buffer<unsigned> k(nullptr, nullptr);
    rocprim::double_buffer<int> v(nullptr, nullptr);
    if (rocprim::radix_sort_pairs(nullptr, bytes, k, v, count, 0, bits) != hipSuccess) return 0;
    return bytes;
}

OrderScratch carve(void* base, int count, std::size_t sortBytes) {
    char* p = static_cast<char*>(base);
    const std::size_t a = alignUp(static_cast<std::size_t>(count) * 4);
    OrderScratch o;
    o.key = reinterpret_cast<unsigned*>(p);
    o.keyAlt = reinterpret_cast<unsigned*>(p + a);
    o.slot = reinterpret_cast<int*>(p + 2 * a);
    o.slotAlt = reinterpret_cast<int*>(p + 3 * a);
    o.sortTmp = p + 4 * a;
    o.sortBytes = sortBytes;
    return o;
}

}  // namespace

std::size_t particleOrderScratchBytes(int count, const Dims& d) {
    if (count <= 0 || d.cells() <= 0) return 0;
    const std::size_t t = sortTempBytes(count, keyBits(d));
    return t ? 4 * alignUp(static_cast<std::size_t>(count) * 4) + alignUp(t) : 0;
}

int launchParticleAdvect(float* pos4, int count, const float* u, const float* v, const float* w,
                         const std::uint8_t* kind, const Dims& d, const ParticleParams& prm, float* vel4,
                         void* orderScratch, std::size_t orderBytes, void* stream) {
    if (!pos4 || count <= 0 || !u || !v || !w || d.cells() <= 0) return static_cast<int>(hipErrorInvalidValue);
    float4* pos = reinterpret_cast<float4*>(pos4);
    const hipStream_t st = static_cast<hipStream_t>(stream);
    const unsigned grid = static_cast<unsigned>((count + 255) / 256);
    if (!vel4) {
        advectKernel<VelSoA><<<grid, 256, 0, st>>>(pos, count, VelSoA{u, v, w, kind}, kind, d, prm, nullptr);
        return static_cast<int>(hipGetLastError());
    }
    float4* vel = reinterpret_cast<float4*>(vel4);
    const long long n = d.cells();
    packVelocityKernel<<<static_cast<unsigned>((n + 255) / 256), 256, 0, st>>>(u, v, w, kind, vel, n);
    const int* order = nullptr;
    if (orderScratch) {
        const int bits = keyBits(d);
        const std::size_t tb = sortTempBytes(count, bits);
        if (!tb || orderBytes < 4 * alignUp(static_cast<std::size_t>(count) * 4) + tb)
            return static_cast<int>(hipErrorInvalidValue);
        OrderScratch o = carve(orderScratch, count, tb);
        particleKeyKernel<<<grid, 256, 0, st>>>(pos, count, d, o.key, o.slot);
        rocprim::double_buffer<unsigned> keys(o.key, o.keyAlt);
        rocprim::double_buffer<int> slots(o.slot, o.slotAlt);
        std::size_t bytes = o.sortBytes;
        if (rocprim::radix_sort_pairs(o.sortTmp, bytes, keys, slots, count, 0, bits, st) != hipSuccess)
            return static_cast<int>(hipErrorUnknown);
        order = slots.current();
    }
    const unsigned g = order ? xcdGrid(grid) : grid;
    advectKernel<VelPacked><<<g, 256, 0, st>>>(pos, count, VelPacked{vel}, kind, d, prm, order);
    return static_cast<int>(hipGetLastError());
}

}  // namespace fc
