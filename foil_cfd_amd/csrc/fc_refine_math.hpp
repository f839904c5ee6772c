// Refinement coupling arithmetic shared by the refinement kernels and the
// fine step that writes restriction records (reference lbm_refine.cu:
// 211-226 for the capped moments, :274-429 for the child average). Every
// function turns floating-point contraction off for its own body, so the
// records come out bit-identical in either translation unit.
#pragma once

#include "fc_kcommon.hpp"

namespace fc {
namespace {

constexpr int kRestrictBand = 2;  // coarse cells skipped at the box edge (lbm_refine.cuh:77)
constexpr float kRefineMaxSpeed = 0.30f;
constexpr int kRecordPlanes = 6;  // float4 planes of one restriction record

// Workgroups are dealt to the XCDs round-robin. A grid of xcdGrid(n)
// workgroups maps to n logical blocks through xcdBlock so that each XCD
// takes one contiguous range of them (neighbouring blocks share an L2);
// blocks at or past n have no work.
inline unsigned xcdGrid(long long n) { return static_cast<unsigned>((n + kXcds - 1) / kXcds * kXcds); }

__device__ inline long long xcdBlock(long long n) {
    const long long chunk = (n + kXcds - 1) / kXcds;
    const long long hb = blockIdx.x;
    return (hb % kXcds) * chunk + hb / kXcds;
}

struct Mom {
    float rho, ux, uy, uz;
};

// Capped moments of a deviation set (density floor 0.05, |u| cap 0.30, as the
// collision; lbm_refine.cu:211-226).
__device__ inline Mom momentsOf(const float h[kQ]) {
#pragma clang fp contract(off)
    float r = 1.0f, jx = 0, jy = 0, jz = 0;
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        r += h[q];
        jx += kCx[q] * h[q];
        jy += kCy[q] * h[q];
        jz += kCz[q] * h[q];
    }
    r = fmaxf(r, 0.05f);
    const float inv = 1.0f / r;
    Mom mo{r, jx * inv, jy * inv, jz * inv};
    const float u2 = mo.ux * mo.ux + mo.uy * mo.uy + mo.uz * mo.uz;
    if (u2 > kRefineMaxSpeed * kRefineMaxSpeed) {
        const float sc = kRefineMaxSpeed * frsqrt(u2);
        mo.ux *= sc;
        mo.uy *= sc;
        mo.uz *= sc;
    }
    return mo;
}

// feq_q(rho, u) - w_q.
__device__ inline float heq(int q, const Mom& mo) {
#pragma clang fp contract(off)
    const float usq = 1.5f * (mo.ux * mo.ux + mo.uy * mo.uy + mo.uz * mo.uz);
    const float cu = 3.0f * (kCx[q] * mo.ux + kCy[q] * mo.uy + kCz[q] * mo.uz);
    return wOf<float>(q) * (mo.rho - 1.0f + mo.rho * (cu + 0.5f * cu * cu - usq));
}

// Restriction record of a coarse cell from the summed deviations h of its nf
// non-solid children: rec[0..18] = fneq of the child average, rec[19] = 1
// (valid), rec[20..23] = its capped moments; all zero when nf = 0.
__device__ inline void childRecord(float h[kQ], int nf, float rec[4 * kRecordPlanes]) {
#pragma clang fp contract(off)
    if (nf == 0) {
#pragma unroll
        for (int k = 0; k < 4 * kRecordPlanes; ++k) rec[k] = 0.0f;
        return;
    }
    const float invN = 1.0f / static_cast<float>(nf);
#pragma unroll
    for (int q = 0; q < kQ; ++q) h[q] *= invN;
    const Mom mo = momentsOf(h);
#pragma unroll
    for (int q = 0; q < kQ; ++q) rec[q] = h[q] - heq(q, mo);
    rec[19] = 1.0f;
    rec[20] = mo.rho;
    rec[21] = mo.ux;
    rec[22] = mo.uy;
    rec[23] = mo.uz;
}

}  // namespace
}  // namespace fc
