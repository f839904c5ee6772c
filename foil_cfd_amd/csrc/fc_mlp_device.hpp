// The learned friction-velocity network u_tau = F(y_s, |u_t|, nu): a
// 3 -> 32 -> 32 -> 1 tanh MLP over normalised log-features, predicting
// log(u_tau / |u_t|). Weights are BF16 (FP32 accumulate); the normalisation
// and the training envelope (per-feature and Re_y = y_s |u_t| / nu bounds)
// ride with the weights so the wall-model kernel can decide per cell whether
// to trust the network or fall back to the Newton solve.
#pragma once

#include <hip/hip_runtime.h>

#include "fc_kernels.hpp"

namespace fc {

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

__device__ inline bool mlpInEnvelope(const MlpParams& m, float ys, float ut, float nu) {
    float f[4];
    mlpFeatures(ys, ut, nu, f);
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (!(f[i] >= m.envLo[i] && f[i] <= m.envHi[i])) return false;
    return true;
}

// Per-thread FP32 inference (BF16 weights); used by the wall-model kernel.
__device__ inline float mlpUTau(const MlpParams& m, float ys, float ut, float nu) {
    float f[4];
    mlpFeatures(ys, ut, nu, f);
    float x[kMlpIn];
#pragma unroll
    for (int i = 0; i < kMlpIn; ++i) x[i] = (f[i] - m.inMean[i]) * m.inInvStd[i];
    float h1[kMlpHidden];
#pragma unroll
    for (int j = 0; j < kMlpHidden; ++j) {
        float acc = m.b1[j];
#pragma unroll
        for (int i = 0; i < kMlpIn; ++i) acc += bf16ToFloat(m.w1[j * kMlpIn + i]) * x[i];
        h1[j] = tanhf(acc);
    }
    float out = m.b3;
#pragma unroll 4
    for (int j = 0; j < kMlpHidden; ++j) {
        float acc = m.b2[j];
#pragma unroll
        for (int i = 0; i < kMlpHidden; ++i) acc += bf16ToFloat(m.w2[j * kMlpHidden + i]) * h1[i];
        out += bf16ToFloat(m.w3[j]) * tanhf(acc);
    }
    return ut * __expf(out * m.outStd + m.outMean);
}

}  // namespace fc
