// The learned friction-velocity network u_tau = F(y_s, |u_t|, nu): a
// 3 -> 32 -> 32 -> 1 tanh MLP over normalised log-features, predicting
// log(u_tau / |u_t|). Weights are BF16 (FP32 accumulate); the normalisation
// and the training envelope (per-feature and Re_y = y_s |u_t| / nu bounds)
// ride with the weights so the wall-model kernel can decide per cell whether
// to trust the network or fall back to the Newton solve.
//
// Inference is wavefront-cooperative: the 64 samples of a wave form two
// 32-sample tiles, and the 32 x 32 hidden layer of each tile is one chain of
// four v_mfma_f32_32x32x8_bf16 (Z2^T = W2 H1^T, FP32 accumulate). The first
// layer (K = 3) and the output layer run on the vector ALUs. The training
// kernels (fc_mlp.hip) use the same operand roundings: BF16 weights, BF16
// first-layer activations into the matrix cores, FP32 biases and FP32
// second-layer activations.
#pragma once

#include <hip/hip_runtime.h>

#include "fc_kernels.hpp"

namespace fc {

typedef short MfmaBf16x4 __attribute__((ext_vector_type(4)));
typedef float MfmaF32x16 __attribute__((ext_vector_type(16)));

// Log-features of one wall sample: (ln y_s, ln u_t, ln nu, ln Re_y).
__host__ __device__ inline void mlpFeatures(float ys, float ut, float nu, float f[4]) {
    f[0] = logf(ys);
    f[1] = logf(ut);
    f[2] = logf(nu);
    f[3] = f[0] + f[1] - f[2];
}

__device__ inline float bf16ToFloat(std::uint16_t b) {
    return __uint_as_float(static_cast<unsigned>(b) << 16);
}

// Round-to-nearest-even FP32 -> BF16 bits (inputs are finite activations).
__device__ inline short bf16Bits(float x) { return __builtin_bit_cast(short, static_cast<__bf16>(x)); }

// tanh(x) = 1 - 2 / (e^{2x} + 1) on the hardware exp2 / reciprocal: absolute
// error ~1e-7, saturates cleanly to +-1 (e^{2x} -> inf or 0). Training and
// inference use the same activation.
__device__ inline float mlpTanh(float x) { return 1.0f - 2.0f / (__expf(2.0f * x) + 1.0f); }

// MFMA 32x32x8 lane layout (wave64), lane l, half h = l / 32:
//   A (32 x 8): A[l % 32][4h + e], e = 0..3;  B (8 x 32): B[4h + e][l % 32];
//   D (32 x 32): register r holds D[8 (r / 4) + 4h + (r % 4)][l % 32].
__device__ inline int mfmaRow(int r, int h) { return 8 * (r >> 2) + 4 * h + (r & 3); }

__device__ inline bool mlpInEnvelope(const MlpParams& m, float ys, float ut, float nu) {
    float f[4];
    mlpFeatures(ys, ut, nu, f);
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (!(f[i] >= m.envLo[i] && f[i] <= m.envHi[i])) return false;
    return true;
}

// Network output (normalised ln(u_tau/u_t)) of every lane's sample. All 64
// lanes of the wave must call it together (full execution mask); a lane with
// use == false contributes a zero feature vector and ignores its result.
__device__ inline float mlpWaveOutput(const MlpParams& m, bool use, float ys, float ut, float nu) {
    const int l = static_cast<int>(threadIdx.x & 63u), s = l & 31, h = l >> 5;
    float x[kMlpIn] = {0.0f, 0.0f, 0.0f};
    if (use) {
        float f[4];
        mlpFeatures(ys, ut, nu, f);
#pragma unroll
        for (int i = 0; i < kMlpIn; ++i) x[i] = (f[i] - m.inMean[i]) * m.inInvStd[i];
    }
    // A operand: W2[j = s][k], this lane's K slices (BF16 bits as stored).
    MfmaBf16x4 aw[4];
#pragma unroll
    for (int kb = 0; kb < 4; ++kb)
#pragma unroll
        for (int e = 0; e < 4; ++e) aw[kb][e] = static_cast<short>(m.w2[s * kMlpHidden + kb * 8 + 4 * h + e]);
    float mine = 0.0f;
#pragma unroll 1
    for (int t = 0; t < 2; ++t) {
        const int src = 32 * t + s;  // sample of column s in tile t
        const float x0 = __shfl(x[0], src, 64), x1 = __shfl(x[1], src, 64), x2 = __shfl(x[2], src, 64);
        MfmaBf16x4 b[4];
#pragma unroll
        for (int kb = 0; kb < 4; ++kb)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int k = kb * 8 + 4 * h + e;
                const float pre = m.b1[k] + bf16ToFloat(m.w1[3 * k]) * x0 + bf16ToFloat(m.w1[3 * k + 1]) * x1 +
                                  bf16ToFloat(m.w1[3 * k + 2]) * x2;
                b[kb][e] = bf16Bits(mlpTanh(pre));
            }
        MfmaF32x16 acc = {};
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(aw[kb], b[kb], acc, 0, 0, 0);
        float o = 0.0f;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int j = mfmaRow(r, h);
            o += bf16ToFloat(m.w3[j]) * mlpTanh(acc[r] + m.b2[j]);
        }
        o += __shfl_xor(o, 32, 64);
        if (t == h) mine = o;  // lane l's own sample is column l % 32 of tile l / 32
    }
    return m.b3 + mine;
}

// Friction velocity from the network for every lane (wave-cooperative).
__device__ inline float mlpWaveUTau(const MlpParams& m, bool use, float ys, float ut, float nu) {
    const float y = mlpWaveOutput(m, use, ys, ut, nu);
    return ut * __expf(y * m.outStd + m.outMean);
}

}  // namespace fc
