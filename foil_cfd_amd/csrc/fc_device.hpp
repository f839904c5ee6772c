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
};

struct StoreF64 {
    using T = double;
    using C = double;
    static constexpr double K = 1.0;
    __device__ static inline C load(const T* p) { return *p; }
    __device__ static inline void store(T* p, C v) { *p = v; }
};

// Fast reciprocal / square root for the FP32 compute path (1 ulp hardware
// ops); the FP64 path keeps correctly rounded IEEE arithmetic.
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
