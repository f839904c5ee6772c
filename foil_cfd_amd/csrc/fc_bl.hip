// Boundary-layer readouts on the live macro fields (reference
// LBMSolver::extractSuctionDelta99 / separationOnsetXc /
// downloadCrossflowPlane, src/sim/lbm_solver.cpp:1511-1636, and the
// streamwise-circulation band integral of src/geom/vg_audit.cpp:27-101).
// The reference downloads whole mid-span planes to the host and scans them
// there; here each probe reads only the cells it needs on the device: one
// wavefront per delta99 station, one lane per column for the separation
// search, one lane per crossflow cell for the circulation band, so a readout
// beside a running sweep moves a few kilobytes instead of two planes.
#include <hip/hip_runtime.h>

#include <climits>

#include "fc_kcommon.hpp"

#pragma clang fp contract(off)

namespace fc {
namespace {

// |(a, b)| rounded to FP32 from the exact double sum of squares: a and b are
// FP32, so the squares are exact in FP64; this is glibc's hypotf.
__device__ inline float speedOf(float a, float b) {
    const double da = a, db = b;
    return static_cast<float>(__builtin_sqrt(da * da + db * db));
}

__device__ inline float waveMaxBcast(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, kWave));
    return v;
}

__device__ inline int waveMinBcast(int v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, kWave));
    return v;
}

// One wavefront per station: pass 1 the edge speed (ray maximum), pass 2 the
// first height reaching 0.99 of it; plus the near-wall streamwise velocity.
__global__ void __launch_bounds__(kWave) delta99Kernel(const float* __restrict__ u,
                                                       const float* __restrict__ v, int nx,
                                                       long long planeOff, const BlStation* st,
                                                       int count, BlProbe* out) {
    const int i = blockIdx.x;
    if (i >= count) return;
    const BlStation s = st[i];
    const int lane = threadIdx.x;
    const float* pu = u + planeOff + s.ix;
    const float* pv = v + planeOff + s.ix;
    float ue = 0.0f;
    for (int y = s.ys + 1 + lane; y <= s.yTop; y += kWave) {
        const long long o = static_cast<long long>(nx) * y;
        ue = fmaxf(ue, speedOf(pu[o], pv[o]));
    }
    ue = waveMaxBcast(ue);
    const float thr = 0.99f * ue;
    int first = INT_MAX;
    for (int y = s.ys + 1 + lane; y <= s.yTop; y += kWave) {
        const long long o = static_cast<long long>(nx) * y;
        if (speedOf(pu[o], pv[o]) >= thr) {
            first = y;
            break;
        }
    }
    first = waveMinBcast(first);
    if (lane == 0) {
        BlProbe r;
        r.ue = ue;
        r.firstY = first == INT_MAX ? -1 : first;
        r.uNear = pu[static_cast<long long>(nx) * (s.ys + 1)];
        r.pad = 0;
        out[i] = r;
    }
}

// Separation onset: the first column i of [ixStart, ixEnd - 2] where columns
// i, i+1, i+2 all have a surface and reversed flow two cells above it.
__device__ inline bool reversedAt(const float* u, const int* topY, int nx, int ny, long long planeOff,
                                  int ix, float uRev) {
    const int ys = topY[ix];
    if (ys < 0 || ys + 2 >= ny) return false;
    return u[planeOff + ix + static_cast<long long>(nx) * (ys + 2)] < uRev;
}

__global__ void __launch_bounds__(kBlock) separationKernel(const float* __restrict__ u,
                                                           const int* __restrict__ topY, int nx, int ny,
                                                           long long planeOff, int ixStart, int ixEnd,
                                                           float uRev, int* outMin) {
    const int ix = ixStart + static_cast<int>(blockIdx.x * kBlock + threadIdx.x);
    if (ix + 2 > ixEnd) return;
    if (reversedAt(u, topY, nx, ny, planeOff, ix, uRev) && reversedAt(u, topY, nx, ny, planeOff, ix + 1, uRev) &&
        reversedAt(u, topY, nx, ny, planeOff, ix + 2, uRev))
        atomicMin(outMin, ix);
}

__global__ void __launch_bounds__(kBlock) minInitKernel(int* p) {
    if (threadIdx.x == 0) *p = INT_MAX;
}

// The (v, w) crossflow plane at column ix, packed y + ny * z.
__global__ void __launch_bounds__(kBlock) crossflowKernel(const float* __restrict__ v,
                                                          const float* __restrict__ w, int nx, int ny, int nz,
                                                          int ix, float* ov, float* ow) {
    const long long k = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (k >= static_cast<long long>(ny) * nz) return;
    const long long c = ix + static_cast<long long>(nx) * k;
    ov[k] = v[c];
    ow[k] = w[c];
}

// Streamwise vorticity omega_x = dw/dy - dv/dz over the band y in
// [yMin, yMax] (all z, periodic) of column ix, central differences with the
// in-band one-sided stencil on the band's first and last row (reference
// integrateBandCirculation, vg_audit.cpp:62-101). omega in FP32 like the
// reference; the four signed sums (zone +, zone -, ambient +, ambient -) in
// FP64, block partials then a fixed-order final sum.
__global__ void __launch_bounds__(kBlock) bandKernel(const float* __restrict__ v, const float* __restrict__ w,
                                                     int nx, int ny, int nz, int ix, int yMin, int yMax,
                                                     int z0, int z1, double* partials) {
    const int rows = yMax - yMin + 1;
    const long long k = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    double acc[4] = {0, 0, 0, 0};
    if (k < static_cast<long long>(rows) * nz) {
        const int z = static_cast<int>(k / rows);
        const int y = yMin + static_cast<int>(k - static_cast<long long>(z) * rows);
        const int zm = z == 0 ? nz - 1 : z - 1;
        const int zp = z == nz - 1 ? 0 : z + 1;
        auto at = [&](const float* a, int yy, int zz) {
            return a[ix + static_cast<long long>(nx) * (yy + static_cast<long long>(ny) * zz)];
        };
        float dwdy;
        if (y == yMin)
            dwdy = at(w, y + 1, z) - at(w, y, z);
        else if (y == yMax)
            dwdy = at(w, y, z) - at(w, y - 1, z);
        else
            dwdy = 0.5f * (at(w, y + 1, z) - at(w, y - 1, z));
        const float dvdz = 0.5f * (at(v, y, zp) - at(v, y, zm));
        const float om = dwdy - dvdz;
        const int slot = (z >= z0 && z <= z1) ? 0 : 2;
        acc[slot + (om > 0.0f ? 0 : 1)] = static_cast<double>(om);
    }
    blockSum<4>(acc, partials + 4 * blockIdx.x);
}

__global__ void __launch_bounds__(kBlock) bandFinalKernel(const double* partials, int slots, double* out) {
    double acc[4] = {0, 0, 0, 0};
    for (int s = threadIdx.x; s < slots; s += kBlock)
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[i] += partials[4 * s + i];
    blockSum<4>(acc, out);
}

int lastErr() { return static_cast<int>(hipGetLastError()); }

}  // namespace

int launchDelta99(const float* u, const float* v, int nx, long long planeOff, const BlStation* stations, int count,
                  BlProbe* out, void* stream) {
    if (count <= 0) return 0;
    delta99Kernel<<<count, kWave, 0, static_cast<hipStream_t>(stream)>>>(u, v, nx, planeOff, stations, count, out);
    return lastErr();
}

int launchSeparationOnset(const float* u, const int* topY, int nx, int ny, long long planeOff, int ixStart,
                          int ixEnd, float uRev, int* outMin, void* stream) {
    const auto st = static_cast<hipStream_t>(stream);
    minInitKernel<<<1, kWave, 0, st>>>(outMin);
    const int cols = ixEnd - 1 - ixStart;  // candidates ixStart .. ixEnd - 2
    if (cols > 0)
        separationKernel<<<(cols + kBlock - 1) / kBlock, kBlock, 0, st>>>(u, topY, nx, ny, planeOff, ixStart, ixEnd,
                                                                           uRev, outMin);
    return lastErr();
}

int launchCrossflowPlane(const float* v, const float* w, const Dims& d, int ix, float* outV, float* outW,
                         void* stream) {
    const long long n = static_cast<long long>(d.ny) * d.nz;
    const unsigned blocks = static_cast<unsigned>((n + kBlock - 1) / kBlock);
    crossflowKernel<<<blocks, kBlock, 0, static_cast<hipStream_t>(stream)>>>(v, w, d.nx, d.ny, d.nz, ix, outV, outW);
    return lastErr();
}

int bandPartialSlots(const Dims& d) { return (d.ny * d.nz + kBlock - 1) / kBlock + 1; }

int launchBandCirculation(const float* v, const float* w, const Dims& d, int ix, int yMin, int yMax, int z0, int z1,
                          double* partials, int slots, double* out4, void* stream) {
    const auto st = static_cast<hipStream_t>(stream);
    const long long n = static_cast<long long>(yMax - yMin + 1) * d.nz;
    const int blocks = static_cast<int>((n + kBlock - 1) / kBlock);
    if (n <= 0 || blocks > slots) return static_cast<int>(hipErrorInvalidValue);
    bandKernel<<<blocks, kBlock, 0, st>>>(v, w, d.nx, d.ny, d.nz, ix, yMin, yMax, z0, z1, partials);
    bandFinalKernel<<<1, kBlock, 0, st>>>(partials, blocks, out4);
    return lastErr();
}

}  // namespace fc
