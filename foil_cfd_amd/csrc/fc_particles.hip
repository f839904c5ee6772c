// Tracer particles for gfx950: lowbias32 seeding and RK2 (midpoint)
// advection through the trilinearly sampled macroscopic velocity field, with
// outflow / out-of-domain / solid-entry / lifetime respawn at the inlet plane
// or in the fluid volume. Same rules, constants and operation order as the
// reference (src/render/particles.cu:48-376); floating-point contraction is
// switched off so every particle is bit-identical to the CPU restatement.
#include <hip/hip_runtime.h>

#include "fc_kcommon.hpp"

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
    __device__ V3 at(long long c) const { return V3{u[c], v[c], w[c]}; }
};
struct VelPacked {
    const float4* vel;
    __device__ V3 at(long long c) const {
        const float4 t = vel[c];
        return V3{t.x, t.y, t.z};
    }
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

// (u, v, w) of every cell into one 16-byte vector (read once per cell).
__global__ void __launch_bounds__(256) packVelocityKernel(const float* u, const float* v, const float* w,
                                                          float4* vel, long long n) {
    const long long c = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c < n) vel[c] = make_float4(u[c], v[c], w[c], 0.0f);
}

template <class VS>
__global__ void __launch_bounds__(256) advectKernel(float4* pos, int count, VS vs, const std::uint8_t* kind, Dims d,
                                                    ParticleParams prm) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
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
    const bool solid = !outflow && !oob && kind && isSolid(kind, cellAt(d, cx, cy, cz));
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

int launchParticleAdvect(float* pos4, int count, const float* u, const float* v, const float* w,
                         const std::uint8_t* kind, const Dims& d, const ParticleParams& prm, float* vel4,
                         void* stream) {
    if (!pos4 || count <= 0 || !u || !v || !w || d.cells() <= 0) return static_cast<int>(hipErrorInvalidValue);
    float4* pos = reinterpret_cast<float4*>(pos4);
    const hipStream_t st = static_cast<hipStream_t>(stream);
    if (vel4) {
        float4* vel = reinterpret_cast<float4*>(vel4);
        const long long n = d.cells();
        packVelocityKernel<<<static_cast<unsigned>((n + 255) / 256), 256, 0, st>>>(u, v, w, vel, n);
        advectKernel<VelPacked><<<(count + 255) / 256, 256, 0, st>>>(pos, count, VelPacked{vel}, kind, d, prm);
    } else {
        advectKernel<VelSoA><<<(count + 255) / 256, 256, 0, st>>>(pos, count, VelSoA{u, v, w}, kind, d, prm);
    }
    return static_cast<int>(hipGetLastError());
}

}  // namespace fc
