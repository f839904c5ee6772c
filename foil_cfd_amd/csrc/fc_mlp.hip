// Training kernels of the learned friction-velocity network (no reference
// counterpart: the reference solves Reichardt's law by Newton only).
//
// Network: x (3 normalised log-features) -> h1 = tanh(W1 x + b1) (32)
//          -> h2 = tanh(W2 h1 + b2) (32) -> y = w3 . h2 + b3,
// trained on mean squared error against the normalised log(u_tau / u_t) of
// harvested Newton solutions. FP32 master parameters, packed
//   W1[32][3] @0, b1[32] @96, W2[32][32] @128 (row = output unit),
//   b2[32] @1152, w3[32] @1184, b3 @1216   (1217 floats).
// The forward pass uses BF16-rounded weights (as the inference kernel does)
// and FP32 biases. The 32 x 32 hidden layer and the two backward GEMMs
// (dW2 = dZ2^T H1 and dH1 = dZ2 W2) run on the matrix cores as
// v_mfma_f32_32x32x8_bf16 tiles with FP32 accumulate: one wavefront per
// 32-sample tile, layout changes through LDS. Per-tile gradient partials are
// reduced in a fixed order (deterministic), then an Adam step updates the
// master parameters.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "fc_mlp_device.hpp"

namespace fc {
namespace {

constexpr int kH = 32;
constexpr int oW1 = 0, oB1 = 96, oW2 = 128, oB2 = 1152, oW3 = 1184, oB3 = 1216;
constexpr int kNP = 1217;
constexpr int kTile = 32;  // samples per wavefront

using v4s = MfmaBf16x4;
using v16f = MfmaF32x16;

__device__ inline float bfRound(float x) { return static_cast<float>(static_cast<__bf16>(x)); }
__device__ inline short bfBits(float x) { return bf16Bits(x); }
__device__ inline int dRow(int r, int h) { return mfmaRow(r, h); }

__global__ void __launch_bounds__(64) mlpGradKernel(const float* __restrict__ X, long long n,
                                                   const float* __restrict__ T, const int* __restrict__ idx,
                                                   int batch, const float* __restrict__ P,
                                                   float* __restrict__ part, float* __restrict__ lossPart) {
    __shared__ float sW[kNP];
    __shared__ float sH1[kTile][kH + 1];
    __shared__ float sH2[kTile][kH + 1];
    __shared__ float sDZ[kTile][kH + 1];
    __shared__ float sX[kTile][3];
    __shared__ float sDy[kTile];
    const int l = threadIdx.x, s = l & 31, h = l >> 5;
    const int tile = blockIdx.x;
    for (int i = l; i < kNP; i += 64) sW[i] = P[i];
    const long long g = idx[tile * kTile + s];
    const float x0 = X[g], x1 = X[n + g], x2 = X[2 * n + g];
    const float t = T[g];
    if (h == 0) {
        sX[s][0] = x0;
        sX[s][1] = x1;
        sX[s][2] = x2;
    }
    __syncthreads();

    // ---- forward: H1 (VALU, K = 3), in the A layout ----------------------
    v4s a[4];
#pragma unroll
    for (int kb = 0; kb < 4; ++kb)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int k = kb * 8 + 4 * h + e;
            const float pre = sW[oB1 + k] + bfRound(sW[oW1 + 3 * k]) * x0 + bfRound(sW[oW1 + 3 * k + 1]) * x1 +
                              bfRound(sW[oW1 + 3 * k + 2]) * x2;
            const float hv = mlpTanh(pre);
            sH1[s][k] = hv;
            a[kb][e] = bfBits(hv);
        }
    // ---- Z2 = H1 W2^T on the matrix cores (M = sample, N = out, K = in) ---
    v16f acc = {};
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
        v4s b;
#pragma unroll
        for (int e = 0; e < 4; ++e) b[e] = bfBits(sW[oW2 + s * kH + kb * 8 + 4 * h + e]);
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a[kb], b, acc, 0, 0, 0);
    }
    float h2[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        h2[r] = mlpTanh(acc[r] + sW[oB2 + s]);
        sH2[dRow(r, h)][s] = h2[r];
    }
    __syncthreads();

    // ---- output and loss (lanes of the first half own one sample each) ----
    float e2 = 0.0f;
    if (h == 0) {
        float y = sW[oB3];
#pragma unroll
        for (int j = 0; j < kH; ++j) y += bfRound(sW[oW3 + j]) * sH2[s][j];
        const float e = y - t;
        e2 = e * e;
        sDy[s] = 2.0f * e / static_cast<float>(batch);
    }
    __syncthreads();

    // ---- backward: dZ2 (D layout: column j = s, 16 sample rows) ----------
    const float w3j = bfRound(sW[oW3 + s]);
    float dw3 = 0.0f, db2 = 0.0f;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int i = dRow(r, h);
        const float dy = sDy[i];
        dw3 += dy * h2[r];
        const float dz = dy * w3j * (1.0f - h2[r] * h2[r]);
        db2 += dz;
        sDZ[i][s] = dz;
    }
    dw3 += __shfl_xor(dw3, 32, 64);
    db2 += __shfl_xor(db2, 32, 64);
    __syncthreads();

    // dW2[j][k] = sum_i dZ2[i][j] H1[i][k]   (M = j, N = k, K = i)
    v16f accW = {};
    // dH1[i][k] = sum_j dZ2[i][j] W2[j][k]   (M = i, N = k, K = j)
    v16f accH = {};
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
        v4s aw, bw, ah, bh;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int kk = kb * 8 + 4 * h + e;
            aw[e] = bfBits(sDZ[kk][s]);                // A[j = s][i = kk]
            bw[e] = bfBits(sH1[kk][s]);                // B[i = kk][k = s]
            ah[e] = bfBits(sDZ[s][kk]);                // A[i = s][j = kk]
            bh[e] = bfBits(sW[oW2 + kk * kH + s]);     // B[j = kk][k = s]
        }
        accW = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(aw, bw, accW, 0, 0, 0);
        accH = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(ah, bh, accH, 0, 0, 0);
    }
    // dZ1 = dH1 * (1 - H1^2); per-unit (k = s) sums over this lane's 16 rows
    float db1 = 0.0f, dw1[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int i = dRow(r, h);
        const float hv = sH1[i][s];
        const float dz1 = accH[r] * (1.0f - hv * hv);
        db1 += dz1;
        dw1[0] += dz1 * sX[i][0];
        dw1[1] += dz1 * sX[i][1];
        dw1[2] += dz1 * sX[i][2];
    }
    db1 += __shfl_xor(db1, 32, 64);
#pragma unroll
    for (int c = 0; c < 3; ++c) dw1[c] += __shfl_xor(dw1[c], 32, 64);

    // ---- per-tile partials ------------------------------------------------
    float* out = part + static_cast<long long>(tile) * kNP;
#pragma unroll
    for (int r = 0; r < 16; ++r) out[oW2 + dRow(r, h) * kH + s] = accW[r];
    if (h == 0) {
        out[oW1 + 3 * s] = dw1[0];
        out[oW1 + 3 * s + 1] = dw1[1];
        out[oW1 + 3 * s + 2] = dw1[2];
        out[oB1 + s] = db1;
        out[oB2 + s] = db2;
        out[oW3 + s] = dw3;
    }
    // db3 and the loss: sums over the 32 samples of the first half
    float dyS = h == 0 ? sDy[s] : 0.0f;
#pragma unroll
    for (int off = 16; off > 0; off >>= 1) {
        dyS += __shfl_xor(dyS, off, 64);
        e2 += __shfl_xor(e2, off, 64);
    }
    if (l == 0) {
        out[oB3] = dyS;
        lossPart[tile] = e2;
    }
}

// Fixed-order sum of the per-tile partials: one thread per parameter (and
// one for the loss).
__global__ void __launch_bounds__(256) mlpReduceKernel(const float* __restrict__ part, const float* __restrict__ lossPart,
                                                       int tiles, float* __restrict__ grad, float* __restrict__ loss) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p < kNP) {
        float sum = 0.0f;
        for (int t = 0; t < tiles; ++t) sum += part[static_cast<long long>(t) * kNP + p];
        grad[p] = sum;
    } else if (p == kNP) {
        float sum = 0.0f;
        for (int t = 0; t < tiles; ++t) sum += lossPart[t];
        *loss = sum;
    }
}

// Adam (Kingma & Ba 2015) on the FP32 master parameters.
__global__ void __launch_bounds__(256) mlpAdamKernel(float* __restrict__ P, float* __restrict__ M, float* __restrict__ V,
                                                     const float* __restrict__ grad, float lr, float beta1,
                                                     float beta2, float eps, float c1, float c2) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= kNP) return;
    const float g = grad[p];
    const float m = beta1 * M[p] + (1.0f - beta1) * g;
    const float v = beta2 * V[p] + (1.0f - beta2) * g * g;
    M[p] = m;
    V[p] = v;
    P[p] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

// Minibatch sample indices from a counter-based hash of (seed, step, slot)
// (the lowbias32 finaliser), so a training step needs no host round trip.
__global__ void __launch_bounds__(256) mlpBatchKernel(int* __restrict__ idx, int batch, long long n,
                                                      unsigned seed, unsigned step) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= batch) return;
    unsigned x = seed ^ (step * 0x9E3779B9u) ^ (static_cast<unsigned>(i) * 0x85EBCA6Bu);
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    idx[i] = static_cast<int>(static_cast<unsigned long long>(x) * static_cast<unsigned long long>(n) >> 32);
}

// Standalone batched inference over (y_s, |u_t|, nu) samples (4 floats each;
// the fourth is ignored): the same wave-cooperative routine the wall-model
// kernel runs. out = network u_tau, or -1 outside the training envelope.
__global__ void __launch_bounds__(256) mlpInferKernel(const MlpParams* __restrict__ m, const float* __restrict__ smp,
                                                      int count, float* __restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const bool in = i < count;
    float ys = 1.0f, ut = 1.0f, nu = 1.0f;
    if (in) {
        ys = smp[4 * static_cast<long long>(i)];
        ut = smp[4 * static_cast<long long>(i) + 1];
        nu = smp[4 * static_cast<long long>(i) + 2];
    }
    const bool use = in && ys > 0.0f && ut > 0.0f && nu > 0.0f && mlpInEnvelope(*m, ys, ut, nu);
    const float u = mlpWaveUTau(*m, use, ys, ut, nu);
    if (in) out[i] = use ? u : -1.0f;
}

}  // namespace

int launchMlpBatch(int* idx, int batch, long long n, unsigned seed, unsigned step, void* stream) {
    if (!idx || batch <= 0 || n <= 0 || n > 0x7FFFFFFFLL) return static_cast<int>(hipErrorInvalidValue);
    mlpBatchKernel<<<(batch + 255) / 256, 256, 0, static_cast<hipStream_t>(stream)>>>(idx, batch, n, seed, step);
    return static_cast<int>(hipGetLastError());
}

int launchMlpInfer(const MlpParams* m, const float* samples, int count, float* out, void* stream) {
    if (!m || count < 0 || (count > 0 && (!samples || !out))) return static_cast<int>(hipErrorInvalidValue);
    if (count > 0)
        mlpInferKernel<<<(count + 255) / 256, 256, 0, static_cast<hipStream_t>(stream)>>>(m, samples, count, out);
    return static_cast<int>(hipGetLastError());
}

int launchMlpGradient(const float* X, long long n, const float* T, const int* idx, int batch, const float* P,
                      float* part, float* lossPart, float* grad, float* loss, void* stream) {
    if (!X || !T || !idx || !P || !part || !lossPart || !grad || !loss || batch <= 0 || batch % kTile != 0)
        return static_cast<int>(hipErrorInvalidValue);
    const hipStream_t st = static_cast<hipStream_t>(stream);
    const int tiles = batch / kTile;
    mlpGradKernel<<<tiles, 64, 0, st>>>(X, n, T, idx, batch, P, part, lossPart);
    mlpReduceKernel<<<(kNP + 1 + 255) / 256, 256, 0, st>>>(part, lossPart, tiles, grad, loss);
    return static_cast<int>(hipGetLastError());
}

int launchMlpAdam(float* P, float* M, float* V, const float* grad, float lr, float beta1, float beta2, float eps,
                  int step, void* stream) {
    if (!P || !M || !V || !grad || step < 1) return static_cast<int>(hipErrorInvalidValue);
    const float c1 = 1.0f - powf(beta1, static_cast<float>(step));
    const float c2 = 1.0f - powf(beta2, static_cast<float>(step));
    mlpAdamKernel<<<(kNP + 255) / 256, 256, 0, static_cast<hipStream_t>(stream)>>>(P, M, V, grad, lr, beta1, beta2,
                                                                                  eps, c1, c2);
    return static_cast<int>(hipGetLastError());
}

}  // namespace fc
