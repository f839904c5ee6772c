// iMEM slip-velocity wall function for gfx950 (one thread per listed wall
// cell, run once per stepN batch like the reference's wallModelUpdateKernel,
// src/sim/lbm_wallmodel.cu:76-196): moments at the sampling cell, tangential
// speed, friction velocity from Reichardt's law (Newton, warm-started, or the
// learned u_tau network with the Newton solve as fallback outside its
// envelope), EMA, the closed-form slip velocity that makes the bounce-back
// links carry the modelled shear, low-y+ fade, clamp, FP16 scatter.
#include <hip/hip_runtime.h>

#include "fc_kcommon.hpp"
#include "fc_mlp_device.hpp"

namespace fc {
namespace {

constexpr float kKappa = 0.41f;      // lbm_wallmodel.cuh:27
constexpr float kReichC = 7.8f;      // lbm_wallmodel.cuh:29
constexpr float kFadeY0 = 3.0f;      // lbm_wallmodel.cuh:40
constexpr float kFadeY1 = 6.0f;      // lbm_wallmodel.cuh:41
constexpr float kMaxSimSpeed = 0.30f;

__device__ inline float reichardtUPlus(float yp) {
    return __logf(1.0f + kKappa * yp) / kKappa +
           kReichC * (1.0f - __expf(-yp / 11.0f) - (yp / 11.0f) * __expf(-yp / 3.0f));
}

__device__ inline float reichardtDuPlus(float yp) {
    return 1.0f / (1.0f + kKappa * yp) +
           kReichC * ((1.0f / 11.0f) * __expf(-yp / 11.0f) - (1.0f / 11.0f) * __expf(-yp / 3.0f) +
                      (yp / 33.0f) * __expf(-yp / 3.0f));
}

// Newton on g(u_tau) = u_tau u+(ys u_tau / nu) - u_t (monotone), 6 steps.
__device__ inline float solveUTau(float ut, float ys, float nu, float seed) {
    float utau = seed > 0.0f ? seed : fmaxf(0.04f * ut, sqrtf(nu * ut / ys));
#pragma unroll
    for (int it = 0; it < 6; ++it) {
        const float yp = fmaxf(ys * utau / nu, 1e-8f);
        const float up = reichardtUPlus(yp);
        const float g = utau * up - ut;
        const float dg = up + yp * reichardtDuPlus(yp);
        utau -= g / fmaxf(dg, 1e-6f);
        utau = fmaxf(utau, 0.0f);
    }
    return utau;
}

__device__ inline std::uint16_t toHalfBits(float v) {
    const _Float16 h = static_cast<_Float16>(v);
    std::uint16_t b;
    __builtin_memcpy(&b, &h, 2);
    return b;
}

template <class S, bool MLP>
__global__ void __launch_bounds__(128) wallUpdateKernel(KArgs a, int layout, WallListView wl,
                                                        std::uint16_t* __restrict__ uwx,
                                                        std::uint16_t* __restrict__ uwy,
                                                        std::uint16_t* __restrict__ uwz,
                                                        WallParams p, WallStats* __restrict__ st,
                                                        float* __restrict__ harvest) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool listed = i < wl.count;
    using C = typename S::C;
    const C invK = C(1) / C(S::K);
    const int nxny = a.nx * a.ny;

    // Moments at the sampling cell (same density floor as the collision).
    float rho = 1.0f, ux = 0.0f, uy = 0.0f, uz = 0.0f;
    float nX = 0.0f, nY = 0.0f, nZ = 0.0f, ys = 1.0f;
    if (listed) {
        nX = wl.nx[i];
        nY = wl.ny[i];
        nZ = wl.nz[i];
        ys = wl.ys[i];
        const long long sc = wl.sampleIdx[i];
        const int sz = static_cast<int>(sc / nxny);
        const int srem = static_cast<int>(sc - static_cast<long long>(sz) * nxny);
        const int sy = srem / a.nx, sx = srem - sy * a.nx;
        float jx = 0.0f, jy = 0.0f, jz = 0.0f;
        rho = 0.0f;
#pragma unroll
        for (int q = 0; q < kQ; ++q) {
            const float fq = wOf<float>(q) + static_cast<float>(postH<S>(a, layout, q, sx, sy, sz) * invK);
            rho += fq;
            jx += kCx[q] * fq;
            jy += kCy[q] * fq;
            jz += kCz[q] * fq;
        }
        rho = fmaxf(rho, 0.05f);
        const float inv = 1.0f / rho;
        ux = jx * inv;
        uy = jy * inv;
        uz = jz * inv;
    }
    const float un = ux * nX + uy * nY + uz * nZ;
    const float tx = ux - un * nX, ty = uy - un * nY, tz = uz - un * nZ;
    const float ut = sqrtf(tx * tx + ty * ty + tz * tz);
    const bool active = listed && ut >= p.utCutoff;

    // Learned u_tau: the whole wavefront evaluates the network together
    // (matrix-core tiles); lanes outside the envelope keep the Newton solve.
    bool usedMlp = false;
    float net = 0.0f;
    if constexpr (MLP) {
        usedMlp = active && mlpInEnvelope(*p.mlp, ys, ut, p.nuLat);
        net = mlpWaveUTau(*p.mlp, usedMlp, ys, ut, p.nuLat);
    }
    // Per-lane statistics partials; the wave reduces them before one atomic each.
    float ypSum = 0.0f;
    int ypMaxBits = 0, nActive = 0, nClamped = 0, nMlp = 0;
    if (listed) {
        const long long cell = wl.cellIdx[i];
        if (!active) {  // stagnation / separation: plain bounce-back
            wl.uTau[i] = 0.0f;
            uwx[cell] = uwy[cell] = uwz[cell] = 0;
            if (harvest) {
                float* h = harvest + 4 * static_cast<long long>(i);
                h[0] = ys;
                h[1] = ut;
                h[2] = p.nuLat;
                h[3] = -1.0f;  // no sample
            }
        } else {
            const float itx = tx / ut, ity = ty / ut, itz = tz / ut;
            const float prev = wl.uTau[i];
            const float solved = usedMlp ? net : solveUTau(ut, ys, p.nuLat, prev);
            if (harvest) {  // the Newton solution is the training target and the network's arbiter
                float* h = harvest + 4 * static_cast<long long>(i);
                h[0] = ys;
                h[1] = ut;
                h[2] = p.nuLat;
                h[3] = usedMlp ? solveUTau(ut, ys, p.nuLat, solved) : solved;
            }
            const float utau = prev > 0.0f ? prev + p.emaAlpha * (solved - prev) : solved;
            wl.uTau[i] = utau;

            // Linear iMEM balance over the masked (solid) pull links: F_t = M0 - u_w D.
            float m0 = 0.0f, dcoef = 0.0f;
            const std::uint32_t mask = wl.mask[i];
            const int cz = static_cast<int>(cell / nxny);
            const int crem = static_cast<int>(cell - static_cast<long long>(cz) * nxny);
            const int cy = crem / a.nx, cx = crem - cy * a.nx;
#pragma unroll
            for (int q = 1; q < kQ; ++q) {
                if (!(mask & (1u << q))) continue;
                const int o = kOpp[q];  // outgoing direction into the wall
                const float ct = -(kCx[q] * itx + kCy[q] * ity + kCz[q] * itz);
                const float fo = wOf<float>(o) + static_cast<float>(postH<S>(a, layout, o, cx, cy, cz) * invK);
                m0 += 2.0f * fo * ct;
                dcoef += 6.0f * wOf<float>(q) * ct * ct;
            }
            const float maxN = fmaxf(fabsf(nX), fmaxf(fabsf(nY), fabsf(nZ)));
            const float aw = 1.0f / fmaxf(maxN, 0.1f);
            float uw = 0.0f;
            if (dcoef > 1e-6f) {
                const float target = rho * utau * utau * aw;
                const float sgn = m0 >= 0.0f ? 1.0f : -1.0f;
                uw = (m0 - sgn * target) / dcoef;
            }
            const float ypS = ys * utau / p.nuLat;
            const float fade = fminf(fmaxf((ypS - kFadeY0) / (kFadeY1 - kFadeY0), 0.0f), 1.0f);
            uw *= fade;
            const float cap = fminf(0.5f * ut, 0.5f * kMaxSimSpeed);
            const float uwc = fminf(fmaxf(uw, -cap), cap);
            uwx[cell] = toHalfBits(uwc * itx);
            uwy[cell] = toHalfBits(uwc * ity);
            uwz[cell] = toHalfBits(uwc * itz);

            ypSum = ypS;
            ypMaxBits = __float_as_int(fmaxf(ypS, 0.0f));  // non-negative floats order as ints
            nActive = 1;
            nClamped = uwc != uw ? 1 : 0;
            nMlp = usedMlp ? 1 : 0;
        }
    }
    ypSum = waveSumF(ypSum);
    ypMaxBits = waveMaxI(ypMaxBits);
    nActive = waveSumI(nActive);
    nClamped = waveSumI(nClamped);
    nMlp = waveSumI(nMlp);
    if ((threadIdx.x & (kWave - 1)) == 0 && nActive > 0) {
        atomicAdd(&st->sumYplus, ypSum);
        atomicMax(&st->maxYplusBits, ypMaxBits);
        atomicAdd(&st->activeCells, nActive);
        if (nClamped) atomicAdd(&st->clampedCells, nClamped);
        if (nMlp) atomicAdd(&st->mlpCells, nMlp);
    }
}

__global__ void zeroStatsKernel(WallStats* st) {
    if (threadIdx.x == 0) *st = WallStats{};
}

template <class S>
struct WallOp {
    static int run(const LatticeView& lat, int layout, const WallListView& wl, std::uint16_t* uwx,
                   std::uint16_t* uwy, std::uint16_t* uwz, const WallParams& p, WallStats* st,
                   float* harvest, hipStream_t s) {
        KArgs a = makeArgs(lat);
        zeroStatsKernel<<<1, 64, 0, s>>>(st);
        if (wl.count > 0) {
            const dim3 grid((wl.count + 127) / 128);
            if (p.mode == kWallMlp && p.mlp)
                wallUpdateKernel<S, true><<<grid, 128, 0, s>>>(a, layout, wl, uwx, uwy, uwz, p, st, harvest);
            else
                wallUpdateKernel<S, false><<<grid, 128, 0, s>>>(a, layout, wl, uwx, uwy, uwz, p, st, harvest);
        }
        return static_cast<int>(hipGetLastError());
    }
};

}  // namespace

int launchWallUpdate(const LatticeView& lat, int layout, const WallListView& wl, std::uint16_t* uwx,
                     std::uint16_t* uwy, std::uint16_t* uwz, const WallParams& p, WallStats* stats,
                     float* harvest, void* stream) {
    if (!uwx || !uwy || !uwz || !stats || (wl.count > 0 && (!wl.cellIdx || !wl.uTau)))
        return static_cast<int>(hipErrorInvalidValue);
    const hipStream_t s = static_cast<hipStream_t>(stream);
    switch (lat.storage) {
        case kStoreF32: return WallOp<StoreF32>::run(lat, layout, wl, uwx, uwy, uwz, p, stats, harvest, s);
        case kStoreF16: return WallOp<StoreF16>::run(lat, layout, wl, uwx, uwy, uwz, p, stats, harvest, s);
        case kStoreBF16: return WallOp<StoreBF16>::run(lat, layout, wl, uwx, uwy, uwz, p, stats, harvest, s);
        case kStoreF16C: return WallOp<StoreF16C>::run(lat, layout, wl, uwx, uwy, uwz, p, stats, harvest, s);
        case kStoreF64: return WallOp<StoreF64>::run(lat, layout, wl, uwx, uwy, uwz, p, stats, harvest, s);
        default: return static_cast<int>(hipErrorInvalidValue);
    }
}

}  // namespace fc
