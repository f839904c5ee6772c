// Device-side building blocks shared by every FoilCFD kernel on gfx950:
// D3Q19 constants, population storage formats, the per-cell kind byte and
// the in-place (AA) slot addressing.
//
// Lattice: D3Q19 in the frozen Krueger ordering of the reference
// (src/sim/lbm_core.cuh:20-68). Opposite directions are consecutive
// (1,2)(3,4)...(17,18), which the pair-form collision exploits.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "fc_layout.hpp"

namespace fc {

// ---------------------------------------------------------------------------
// Population storage formats. Every format stores the scaled deviation
// H_q = K * (f_q - w_q) from the rest weights; the collision runs on H in the
// compute type and never reconstructs f. For FP16 the scale K = 2^15 moves
// typical deviations (1e-7 .. 0.3) out of the subnormal range; multiplying by
// a power of two is exact, and K is folded into the equilibrium constants so
// loads and stores are a single conversion each.
// ---------------------------------------------------------------------------
struct StoreF32 {
    using T = float;
    using C = float;
    using Bits = unsigned;
    static constexpr float K = 1.0f;
    __device__ static inline C load(const T* p) { return *p; }
    __device__ static inline void store(T* p, C v) { *p = v; }
    __device__ static inline C fromBits(Bits b) { return __builtin_bit_cast(float, b); }
    __device__ static inline Bits toBits(C v) { return __builtin_bit_cast(Bits, v); }
};

struct StoreF16 {
    using T = _Float16;
    using C = float;
    using Bits = unsigned short;
    static constexpr float K = 32768.0f;
    __device__ static inline C load(const T* p) { return static_cast<float>(*p); }
    __device__ static inline void store(T* p, C v) { *p = static_cast<_Float16>(v); }
    __device__ static inline C fromBits(Bits b) { return static_cast<float>(__builtin_bit_cast(_Float16, b)); }
    __device__ static inline Bits toBits(C v) { return __builtin_bit_cast(Bits, static_cast<_Float16>(v)); }
    // Stochastic rounding when sto is set (random bits r, the top 13 used):
    // the 13 FP32 mantissa bits a normal FP16 drops get r added and are
    // cleared, so the conversion is exact and rounds up with probability
    // equal to the dropped fraction. Values in FP16's subnormal range
    // (|v| < 2^-14, i.e. |f - w| < 2^-29 after the scale) keep round-to-
    // nearest on the rest. Else round to nearest; one conversion either way.
    __device__ static inline Bits toBitsR(C v, bool sto, unsigned r) {
        const unsigned b = __builtin_bit_cast(unsigned, v);
        const bool finite = (b & 0x7F800000u) != 0x7F800000u;
        const float t = __builtin_bit_cast(float, (b + (r >> 19)) & ~0x1FFFu);
        return __builtin_bit_cast(Bits, static_cast<_Float16>(sto && finite ? t : v));
    }
};

struct StoreBF16 {
    using T = __bf16;
    using C = float;
    using Bits = unsigned short;
    static constexpr float K = 1.0f;
    __device__ static inline C load(const T* p) { return static_cast<float>(*p); }
    __device__ static inline void store(T* p, C v) { *p = static_cast<__bf16>(v); }
    __device__ static inline C fromBits(Bits b) { return static_cast<float>(__builtin_bit_cast(__bf16, b)); }
    __device__ static inline Bits toBits(C v) { return __builtin_bit_cast(Bits, static_cast<__bf16>(v)); }
    // Stochastic rounding when sto is set (random bits r, the top 16 used):
    // r is added to the 16 FP32 bits BF16 drops, then they are cut off.
    // Else, and for infinities and NaN, the round-to-nearest conversion.
    __device__ static inline Bits toBitsR(C v, bool sto, unsigned r) {
        const unsigned b = __builtin_bit_cast(unsigned, v);
        const bool finite = (b & 0x7F800000u) != 0x7F800000u;
        return sto && finite ? static_cast<Bits>((b + (r >> 16)) >> 16) : toBits(v);
    }
};

// FP16C: a 16-bit float of the unscaled deviation with 1 sign, 4 exponent
// and 11 mantissa bits (one mantissa bit more than IEEE FP16, whose fifth
// exponent bit the bounded deviations never need). Normals cover
// [2^-14, 2), a subnormal quantum of 2^-25 below; values at or above the
// largest code (1 + 2047/2048) and infinities saturate to it, NaN encodes as
// the largest negative code. Decoding sign-extends the 16 bits (a signed
// 16-bit load), places the 15 magnitude bits under an FP32 exponent/mantissa
// and rescales by 2^112 (3 VALU; FP32 denormals are preserved in these
// kernels, so subnormal codes decode exactly). Encoding rescales by 2^-112,
// clamps, and rounds the FP32 bits to nearest even at bit 12 with the sign
// bit carried along (bits 27-30 are zero after the clamp, so no carry reaches
// it), then packs (7 VALU).
struct StoreF16C {
    using T = unsigned short;
    using C = float;
    using Bits = unsigned short;
    static constexpr float K = 1.0f;
    __device__ static inline C fromBits(Bits b) {
        const unsigned sx = static_cast<unsigned>(static_cast<int>(static_cast<short>(b)));
        const unsigned u = (sx << 12) & 0x87FFF000u;
        return __builtin_bit_cast(float, u) * 0x1p112f;
    }
    __device__ static inline Bits toBits(C v) {
        constexpr float lim = 0x1.FFEp0f * 0x1p-112f;  // code 0x7FFF
        const float t = __builtin_fminf(__builtin_fmaxf(v * 0x1p-112f, -lim), lim);
        const unsigned b = __builtin_bit_cast(unsigned, t);
        const unsigned r = b + 0x7FFu + ((b >> 12) & 1u);
        return static_cast<Bits>(((r >> 12) & 0x7FFFu) | ((r >> 16) & 0x8000u));
    }
    // Stochastic rounding when sto is set (random bits r, the top 12 used):
    // the packing of toBits with r added below bit 12 instead of the
    // round-to-nearest-even offset, so a value rounds up with probability
    // equal to the dropped fraction. NaN (clamped to the largest magnitude
    // by fmin / fmax) stays saturated.
    __device__ static inline Bits toBitsR(C v, bool sto, unsigned r) {
        constexpr float lim = 0x1.FFEp0f * 0x1p-112f;  // code 0x7FFF
        const float t = __builtin_fminf(__builtin_fmaxf(v * 0x1p-112f, -lim), lim);
        const unsigned b = __builtin_bit_cast(unsigned, t);
        const unsigned u = b + (sto ? (r >> 20) : 0x7FFu + ((b >> 12) & 1u));
        return static_cast<Bits>(((u >> 12) & 0x7FFFu) | ((u >> 16) & 0x8000u));
    }
    __device__ static inline C load(const T* p) { return fromBits(*p); }
    __device__ static inline void store(T* p, C v) { *p = toBits(v); }
};

struct StoreF64 {
    using T = double;
    using C = double;
    using Bits = unsigned long long;
    static constexpr double K = 1.0;
    __device__ static inline C fromBits(Bits b) { return __builtin_bit_cast(double, b); }
    __device__ static inline Bits toBits(C v) { return __builtin_bit_cast(Bits, v); }
    __device__ static inline C load(const T* p) { return *p; }
    __device__ static inline void store(T* p, C v) { *p = v; }
};

// Fast reciprocal / square root for the FP32 compute path (1 ulp hardware
// ops); the FP64 path keeps correctly rounded IEEE arithmetic.
// a * b + c with one rounding. The step kernels compile with contraction
// off and place every fused multiply-add explicitly, so each kernel variant
// executes the same IEEE operations and they agree bit for bit.
__device__ inline float fmac(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ inline double fmac(double a, double b, double c) { return __builtin_fma(a, b, c); }

__device__ inline float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline double frcp(double x) { return 1.0 / x; }
__device__ inline float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ inline double fsqrt(double x) { return __builtin_sqrt(x); }
__device__ inline float frsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ inline double frsqrt(double x) { return 1.0 / __builtin_sqrt(x); }

// IEEE half bits -> float (exact); the wall-slip field is stored as half bits
// like the reference (lbm_core.cuh:201-208).
__device__ inline float halfBits(std::uint16_t b) {
    _Float16 h;
    __builtin_memcpy(&h, &b, 2);
    return static_cast<float>(h);
}

template <class T>
__device__ inline T wOf(int q) {
    return q == 0 ? T(1) / T(3) : (q <= 6 ? T(1) / T(18) : T(1) / T(36));
}

}  // namespace fc
