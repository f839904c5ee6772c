// Mesh-sequencing seed: trilinear upsample of a presolve's macroscopic field
// onto a finer grid of the same domain (reference upsampleMacroKernel,
// src/sim/lbm_refine.cu:437-501). One lane per destination cell; source cell
// centres mapped by the grid ratio, x / y clamped to the border sample, z
// periodic. Contraction is off, so the result is the FP32 expression as
// written, bit for bit.
#include <hip/hip_runtime.h>

#include "fc_kcommon.hpp"

#pragma clang fp contract(off)

namespace fc {
namespace {

__global__ void __launch_bounds__(kBlock) upsampleKernel(const float* __restrict__ sRho, const float* __restrict__ sU,
                                                        const float* __restrict__ sV, const float* __restrict__ sW,
                                                        int snx, int sny, int snz, float* __restrict__ dRho,
                                                        float* __restrict__ dU, float* __restrict__ dV,
                                                        float* __restrict__ dW, int dnx, int dny, int dnz) {
    const long long cell = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    const long long dxy = static_cast<long long>(dnx) * dny;
    if (cell >= dxy * dnz) return;
    const int z = static_cast<int>(cell / dxy);
    const long long rem = cell - z * dxy;
    const int y = static_cast<int>(rem / dnx);
    const int x = static_cast<int>(rem - static_cast<long long>(y) * dnx);
    // destination centre in source cell-centre coordinates
    const float gx = (static_cast<float>(x) + 0.5f) * static_cast<float>(snx) / static_cast<float>(dnx) - 0.5f;
    const float gy = (static_cast<float>(y) + 0.5f) * static_cast<float>(sny) / static_cast<float>(dny) - 0.5f;
    const float gz = (static_cast<float>(z) + 0.5f) * static_cast<float>(snz) / static_cast<float>(dnz) - 0.5f;
    int x0 = static_cast<int>(floorf(gx));
    int y0 = static_cast<int>(floorf(gy));
    const int zf = static_cast<int>(floorf(gz));
    float tx = gx - static_cast<float>(x0);
    float ty = gy - static_cast<float>(y0);
    const float tz = gz - static_cast<float>(zf);
    if (x0 < 0) {
        x0 = 0;
        tx = 0.0f;
    }
    if (y0 < 0) {
        y0 = 0;
        ty = 0.0f;
    }
    int x1 = x0 + 1, y1 = y0 + 1;
    if (x1 > snx - 1) {
        x1 = snx - 1;
        tx = 0.0f;
    }
    if (y1 > sny - 1) {
        y1 = sny - 1;
        ty = 0.0f;
    }
    const int z0 = ((zf % snz) + snz) % snz;
    const int z1 = (z0 + 1) % snz;
    const long long sxy = static_cast<long long>(snx) * sny;
    const long long r00 = static_cast<long long>(snx) * y0 + sxy * z0;
    const long long r10 = static_cast<long long>(snx) * y1 + sxy * z0;
    const long long r01 = static_cast<long long>(snx) * y0 + sxy * z1;
    const long long r11 = static_cast<long long>(snx) * y1 + sxy * z1;
    const long long c[8] = {x0 + r00, x1 + r00, x0 + r10, x1 + r10, x0 + r01, x1 + r01, x0 + r11, x1 + r11};
    const float ux = 1 - tx, uy = 1 - ty, uz = 1 - tz;
    const float wgt[8] = {ux * uy * uz, tx * uy * uz, ux * ty * uz, tx * ty * uz,
                          ux * uy * tz, tx * uy * tz, ux * ty * tz, tx * ty * tz};
    auto lerp = [&](const float* s) {
        float acc = wgt[0] * s[c[0]];
#pragma unroll
        for (int k = 1; k < 8; ++k) acc = acc + wgt[k] * s[c[k]];
        return acc;
    };
    dRho[cell] = lerp(sRho);
    dU[cell] = lerp(sU);
    dV[cell] = lerp(sV);
    dW[cell] = lerp(sW);
}

}  // namespace

int launchUpsampleMacro(const float* sRho, const float* sU, const float* sV, const float* sW, const Dims& s,
                        float* dRho, float* dU, float* dV, float* dW, const Dims& d, void* stream) {
    if (s.nx < 1 || s.ny < 1 || s.nz < 1 || d.nx < 1 || d.ny < 1 || d.nz < 1)
        return static_cast<int>(hipErrorInvalidValue);
    const long long n = static_cast<long long>(d.nx) * d.ny * d.nz;
    upsampleKernel<<<static_cast<unsigned>((n + kBlock - 1) / kBlock), kBlock, 0, static_cast<hipStream_t>(stream)>>>(
        sRho, sU, sV, sW, s.nx, s.ny, s.nz, dRho, dU, dV, dW, d.nx, d.ny, d.nz);
    return static_cast<int>(hipGetLastError());
}

}  // namespace fc
