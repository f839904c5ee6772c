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

// ---------------------------------------------------------------------------
// Restriction records in the scratch array (nR records over the restricted
// box). Full format (FP32 / FP64 storage): six float4 planes, plane c at
// scratch4[c nR + ri]: fneq[4c..4c+3] for c < 4, plane 4 fneq[16..18] and the
// valid mark, plane 5 the moments. Compact format (16-bit storages, 64 of
// the 96 bytes): plane 0 the FP32 moments, then three planes of eight IEEE
// halves at ((uint4*)(scratch4 + nR))[c nR + ri]: fneq[0..7], fneq[8..15],
// (fneq[16..18], valid, 0, 0, 0, 0), fneq scaled by 2^14 (normals down to
// 4e-9, saturating at +-4). Their rounding (2^-11 of fneq) is far below the
// 2^-12 of f - w that the 16-bit population storage itself keeps, and the
// restriction filter and the moments read 40 % fewer bytes.
// ---------------------------------------------------------------------------
template <class S>
struct RecordFormat {
    static constexpr bool kCompact = sizeof(typename S::T) == 2;
};
constexpr float kRecScale = 16384.0f;
constexpr float kRecInvScale = 1.0f / 16384.0f;

__device__ inline unsigned short recHalf(float x) {
    const float y = x * kRecScale;
    const float c = __builtin_isnan(y) ? y : fminf(fmaxf(y, -65504.0f), 65504.0f);
    return __builtin_bit_cast(unsigned short, static_cast<_Float16>(c));
}
__device__ inline unsigned recPair(float a, float b) {
    return static_cast<unsigned>(recHalf(a)) | (static_cast<unsigned>(recHalf(b)) << 16);
}
// Scaled value of half i (0..7) of a compact plane (multiply by kRecInvScale).
__device__ inline float recUnpack(const uint4& v, int i) {
    const unsigned w = i < 2 ? v.x : (i < 4 ? v.y : (i < 6 ? v.z : v.w));
    const unsigned short h = static_cast<unsigned short>((i & 1) ? (w >> 16) : (w & 0xFFFFu));
    return static_cast<float>(__builtin_bit_cast(_Float16, h));
}

template <bool COMPACT>
__device__ inline void storeRecord(float4* scratch4, long long nR, long long ri, const float rec[4 * kRecordPlanes]) {
    if (!COMPACT) {
#pragma unroll
        for (int c = 0; c < kRecordPlanes; ++c)
            scratch4[c * nR + ri] = make_float4(rec[4 * c], rec[4 * c + 1], rec[4 * c + 2], rec[4 * c + 3]);
        return;
    }
    scratch4[ri] = make_float4(rec[20], rec[21], rec[22], rec[23]);
    uint4* const h = reinterpret_cast<uint4*>(scratch4 + nR);
    h[ri] = make_uint4(recPair(rec[0], rec[1]), recPair(rec[2], rec[3]), recPair(rec[4], rec[5]),
                       recPair(rec[6], rec[7]));
    h[nR + ri] = make_uint4(recPair(rec[8], rec[9]), recPair(rec[10], rec[11]), recPair(rec[12], rec[13]),
                            recPair(rec[14], rec[15]));
    // the valid mark is stored unscaled: 1.0 -> 0x3C00, 0 -> 0
    const unsigned valid = rec[19] != 0.0f ? 0x3C00u : 0u;
    h[2 * nR + ri] = make_uint4(recPair(rec[16], rec[17]), static_cast<unsigned>(recHalf(rec[18])) | (valid << 16),
                                0u, 0u);
}

}  // namespace
}  // namespace fc
