// Host side of the learned u_tau network: GPU training over harvested wall
// samples and batched inference (include/foilcfd.h, fc_mlp_*). No reference
// counterpart: the reference solves Reichardt's law per wall cell by Newton
// (src/sim/lbm_wallmodel.cu:55-73), which stays the fallback outside the
// training envelope and the arbiter of the network's error.
//
// Data on the GPU: normalised inputs X (SoA [3][n]) and targets T[n], the
// FP32 master parameters with Adam's moments, the minibatch index buffer and
// the per-tile gradient partials. One training step is four launches on the
// trainer's stream (batch draw, gradient, fixed-order reduction, Adam) with
// no host round trip; losses are read back once per call.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/foilcfd.h"
#include "fc_kernels.hpp"

namespace fc {
namespace {

constexpr int kNP = kMlpParamCount;
constexpr int kOW1 = 0, kOB1 = 96, kOW2 = 128, kOB2 = 1152, kOW3 = 1184, kOB3 = 1216;
constexpr int kMaxBatch = 1 << 20;
constexpr double kBeta1 = 0.9, kBeta2 = 0.999, kAdamEps = 1e-8;

template <class T>
bool devAlloc(T** p, std::size_t count) {
    void* raw = nullptr;
    *p = nullptr;
    if (count == 0) return true;
    if (hipMalloc(&raw, count * sizeof(T)) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    *p = static_cast<T*>(raw);
    return true;
}

template <class T>
void devFree(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

// Round to nearest even FP32 -> BF16 bits (finite inputs).
std::uint16_t bf16BitsHost(float x) {
    std::uint32_t u;
    std::memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return static_cast<std::uint16_t>(u >> 16);
}

// lowbias32 integer hash (deterministic weight initialisation).
std::uint32_t mix32(std::uint32_t x) {
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    return x;
}

}  // namespace

struct MlpTrainer {
    hipStream_t stream = nullptr;
    long long n = 0;
    std::uint32_t seed = 0;
    int step = 0;       // Adam steps taken
    unsigned draws = 0;  // minibatches drawn (hash counter)
    float inMean[3] = {}, inInvStd[3] = {}, outMean = 0.0f, outStd = 1.0f;
    float envLo[4] = {}, envHi[4] = {};
    float *dX = nullptr, *dT = nullptr, *dP = nullptr, *dM = nullptr, *dV = nullptr, *dG = nullptr;
    float *dPscratch = nullptr, *dPart = nullptr, *dLossPart = nullptr, *dLoss = nullptr;
    int* dIdx = nullptr;
    int cap = 0;  // batch capacity of dIdx / dPart / dLossPart

    ~MlpTrainer() {
        if (stream) (void)hipStreamSynchronize(stream);
        devFree(dX);
        devFree(dT);
        devFree(dP);
        devFree(dM);
        devFree(dV);
        devFree(dG);
        devFree(dPscratch);
        devFree(dPart);
        devFree(dLossPart);
        devFree(dLoss);
        devFree(dIdx);
        if (stream) (void)hipStreamDestroy(stream);
    }

    bool ensureBatch(int batch) {
        if (batch <= cap) return true;
        devFree(dIdx);
        devFree(dPart);
        devFree(dLossPart);
        cap = 0;
        const int tiles = batch / kMlpTile;
        if (!devAlloc(&dIdx, batch) || !devAlloc(&dPart, static_cast<std::size_t>(tiles) * kNP) ||
            !devAlloc(&dLossPart, tiles))
            return false;
        cap = batch;
        return true;
    }
};

namespace {

// Kept samples -> normalised features / targets, normalisation and envelope.
bool prepare(MlpTrainer& t, const float* smp, long long count, std::vector<float>& X, std::vector<float>& T,
             std::string& err) {
    std::vector<double> f0, f1, f2, tg;
    f0.reserve(count);
    f1.reserve(count);
    f2.reserve(count);
    tg.reserve(count);
    for (long long i = 0; i < count; ++i) {
        const float ys = smp[4 * i], ut = smp[4 * i + 1], nu = smp[4 * i + 2], utau = smp[4 * i + 3];
        if (!(ys > 0.0f && ut > 0.0f && nu > 0.0f && utau > 0.0f) || !std::isfinite(ys) || !std::isfinite(ut) ||
            !std::isfinite(nu) || !std::isfinite(utau))
            continue;
        f0.push_back(std::log(static_cast<double>(ys)));
        f1.push_back(std::log(static_cast<double>(ut)));
        f2.push_back(std::log(static_cast<double>(nu)));
        tg.push_back(std::log(static_cast<double>(utau) / static_cast<double>(ut)));
    }
    const long long n = static_cast<long long>(tg.size());
    if (n < kMlpTile) {
        err = "fewer than 32 usable samples";
        return false;
    }
    if (n > 0x7FFFFFFFLL) {
        err = "more than 2^31-1 samples";
        return false;
    }
    auto stats = [n](const std::vector<double>& v, float& mean, float& invStd) {
        double s = 0.0, s2 = 0.0;
        for (double x : v) s += x;
        const double m = s / n;
        for (double x : v) s2 += (x - m) * (x - m);
        const double sd = std::sqrt(s2 / n);
        mean = static_cast<float>(m);
        invStd = sd > 1e-6 ? static_cast<float>(1.0 / sd) : 1.0f;  // constant feature: centred only
    };
    const std::vector<double>* feats[3] = {&f0, &f1, &f2};
    for (int c = 0; c < 3; ++c) stats(*feats[c], t.inMean[c], t.inInvStd[c]);
    float tInv = 1.0f;
    stats(tg, t.outMean, tInv);
    t.outStd = 1.0f / tInv;
    // Envelope in log-feature space (ln y_s, ln u_t, ln nu, ln Re_y), widened
    // by 1% of its width (at least 1e-3) on each side.
    double lo[4], hi[4];
    for (int c = 0; c < 4; ++c) {
        lo[c] = 1e300;
        hi[c] = -1e300;
    }
    for (long long i = 0; i < n; ++i) {
        const double v[4] = {f0[i], f1[i], f2[i], f0[i] + f1[i] - f2[i]};
        for (int c = 0; c < 4; ++c) {
            lo[c] = std::min(lo[c], v[c]);
            hi[c] = std::max(hi[c], v[c]);
        }
    }
    for (int c = 0; c < 4; ++c) {
        const double pad = std::max(0.01 * (hi[c] - lo[c]), 1e-3);
        t.envLo[c] = static_cast<float>(lo[c] - pad);
        t.envHi[c] = static_cast<float>(hi[c] + pad);
    }
    X.resize(3 * n);
    T.resize(n);
    for (long long i = 0; i < n; ++i) {
        X[i] = static_cast<float>((f0[i] - t.inMean[0]) * t.inInvStd[0]);
        X[n + i] = static_cast<float>((f1[i] - t.inMean[1]) * t.inInvStd[1]);
        X[2 * n + i] = static_cast<float>((f2[i] - t.inMean[2]) * t.inInvStd[2]);
        T[i] = static_cast<float>((tg[i] - t.outMean) * tInv);
    }
    t.n = n;
    return true;
}

// Glorot-uniform weights (Glorot & Bengio 2010), zero biases.
void initParams(std::uint32_t seed, float* P) {
    std::fill(P, P + kNP, 0.0f);
    std::uint32_t c = 0;
    auto uni = [&](double limit) {
        const std::uint32_t r = mix32(seed ^ mix32(++c * 0x9E3779B9u));
        return static_cast<float>(limit * (2.0 * (r * (1.0 / 4294967296.0)) - 1.0));
    };
    for (int i = 0; i < 96; ++i) P[kOW1 + i] = uni(std::sqrt(6.0 / (3 + 32)));
    for (int i = 0; i < 1024; ++i) P[kOW2 + i] = uni(std::sqrt(6.0 / (32 + 32)));
    for (int i = 0; i < 32; ++i) P[kOW3 + i] = uni(std::sqrt(6.0 / (32 + 1)));
}

void setErr(char* err, int errlen, const std::string& msg) {
    if (err && errlen > 0) std::snprintf(err, static_cast<std::size_t>(errlen), "%s", msg.c_str());
}

}  // namespace
}  // namespace fc

struct fc_mlp_trainer {
    fc::MlpTrainer impl;
};

extern "C" {

fc_mlp_trainer* fc_mlp_trainer_create(const float* samples, long long count, uint32_t seed, char* err,
                                      int errlen) {
    if (!samples || count <= 0) {
        fc::setErr(err, errlen, "no samples");
        return nullptr;
    }
    auto* t = new fc_mlp_trainer;
    fc::MlpTrainer& m = t->impl;
    std::vector<float> X, T;
    std::string msg;
    if (!fc::prepare(m, samples, count, X, T, msg)) {
        fc::setErr(err, errlen, msg);
        delete t;
        return nullptr;
    }
    m.seed = seed;
    std::vector<float> P(fc::kNP);
    fc::initParams(seed, P.data());
    const std::vector<float> zeros(fc::kNP, 0.0f);
    const std::size_t pb = fc::kNP * sizeof(float);
    bool ok = hipStreamCreate(&m.stream) == hipSuccess;
    ok = ok && fc::devAlloc(&m.dX, X.size()) && fc::devAlloc(&m.dT, T.size()) && fc::devAlloc(&m.dP, fc::kNP) &&
         fc::devAlloc(&m.dM, fc::kNP) && fc::devAlloc(&m.dV, fc::kNP) && fc::devAlloc(&m.dG, fc::kNP) &&
         fc::devAlloc(&m.dPscratch, fc::kNP) && fc::devAlloc(&m.dLoss, 1);
    ok = ok && hipMemcpyAsync(m.dX, X.data(), X.size() * sizeof(float), hipMemcpyHostToDevice, m.stream) == hipSuccess &&
         hipMemcpyAsync(m.dT, T.data(), T.size() * sizeof(float), hipMemcpyHostToDevice, m.stream) == hipSuccess &&
         hipMemcpyAsync(m.dP, P.data(), pb, hipMemcpyHostToDevice, m.stream) == hipSuccess &&
         hipMemcpyAsync(m.dM, zeros.data(), pb, hipMemcpyHostToDevice, m.stream) == hipSuccess &&
         hipMemcpyAsync(m.dV, zeros.data(), pb, hipMemcpyHostToDevice, m.stream) == hipSuccess &&
         hipStreamSynchronize(m.stream) == hipSuccess;
    if (!ok) {
        (void)hipGetLastError();
        fc::setErr(err, errlen, "device allocation or upload failed");
        delete t;
        return nullptr;
    }
    return t;
}

void fc_mlp_trainer_destroy(fc_mlp_trainer* t) { delete t; }

long long fc_mlp_trainer_samples(fc_mlp_trainer* t) { return t ? t->impl.n : -FC_ERR_INVALID; }

int fc_mlp_trainer_train(fc_mlp_trainer* t, int steps, int batch, float lr, float* losses) {
    if (!t || steps < 0 || batch <= 0 || batch % fc::kMlpTile != 0 || batch > fc::kMaxBatch || !(lr > 0.0f))
        return FC_ERR_INVALID;
    fc::MlpTrainer& m = t->impl;
    if (steps == 0) return FC_OK;
    if (!m.ensureBatch(batch)) return FC_ERR_ALLOC;
    float* dHist = nullptr;
    if (!fc::devAlloc(&dHist, steps)) return FC_ERR_ALLOC;
    int rc = FC_OK;
    for (int s = 0; s < steps && rc == FC_OK; ++s) {
        if (fc::launchMlpBatch(m.dIdx, batch, m.n, m.seed, m.draws++, m.stream) != 0 ||
            fc::launchMlpGradient(m.dX, m.n, m.dT, m.dIdx, batch, m.dP, m.dPart, m.dLossPart, m.dG, dHist + s,
                                  m.stream) != 0 ||
            fc::launchMlpAdam(m.dP, m.dM, m.dV, m.dG, lr, static_cast<float>(fc::kBeta1),
                              static_cast<float>(fc::kBeta2), static_cast<float>(fc::kAdamEps), ++m.step,
                              m.stream) != 0)
            rc = FC_ERR_DEVICE;
    }
    std::vector<float> hist(steps);
    if (rc == FC_OK &&
        (hipMemcpyAsync(hist.data(), dHist, steps * sizeof(float), hipMemcpyDeviceToHost, m.stream) != hipSuccess ||
         hipStreamSynchronize(m.stream) != hipSuccess))
        rc = FC_ERR_DEVICE;
    (void)hipStreamSynchronize(m.stream);
    fc::devFree(dHist);
    if (rc != FC_OK) {
        (void)hipGetLastError();
        return rc;
    }
    if (losses)
        for (int s = 0; s < steps; ++s) losses[s] = hist[s] / static_cast<float>(batch);
    return FC_OK;
}

int fc_mlp_trainer_gradient(fc_mlp_trainer* t, const float* params, const int* idx, int batch, float* grad,
                            float* sse) {
    if (!t || !idx || !grad || batch <= 0 || batch % fc::kMlpTile != 0 || batch > fc::kMaxBatch)
        return FC_ERR_INVALID;
    fc::MlpTrainer& m = t->impl;
    for (int i = 0; i < batch; ++i)
        if (idx[i] < 0 || idx[i] >= m.n) return FC_ERR_INVALID;
    if (!m.ensureBatch(batch)) return FC_ERR_ALLOC;
    const float* dParams = m.dP;
    if (params) {
        if (hipMemcpyAsync(m.dPscratch, params, fc::kNP * sizeof(float), hipMemcpyHostToDevice, m.stream) !=
            hipSuccess)
            return FC_ERR_DEVICE;
        dParams = m.dPscratch;
    }
    float loss = 0.0f;
    if (hipMemcpyAsync(m.dIdx, idx, static_cast<std::size_t>(batch) * sizeof(int), hipMemcpyHostToDevice,
                       m.stream) != hipSuccess ||
        fc::launchMlpGradient(m.dX, m.n, m.dT, m.dIdx, batch, dParams, m.dPart, m.dLossPart, m.dG, m.dLoss,
                              m.stream) != 0 ||
        hipMemcpyAsync(grad, m.dG, fc::kNP * sizeof(float), hipMemcpyDeviceToHost, m.stream) != hipSuccess ||
        hipMemcpyAsync(&loss, m.dLoss, sizeof(float), hipMemcpyDeviceToHost, m.stream) != hipSuccess ||
        hipStreamSynchronize(m.stream) != hipSuccess) {
        (void)hipGetLastError();
        return FC_ERR_DEVICE;
    }
    if (sse) *sse = loss;
    return FC_OK;
}

int fc_mlp_trainer_get_params(fc_mlp_trainer* t, float* params) {
    if (!t || !params) return FC_ERR_INVALID;
    fc::MlpTrainer& m = t->impl;
    if (hipMemcpyAsync(params, m.dP, fc::kNP * sizeof(float), hipMemcpyDeviceToHost, m.stream) != hipSuccess ||
        hipStreamSynchronize(m.stream) != hipSuccess)
        return FC_ERR_DEVICE;
    return FC_OK;
}

int fc_mlp_trainer_set_params(fc_mlp_trainer* t, const float* params) {
    if (!t || !params) return FC_ERR_INVALID;
    fc::MlpTrainer& m = t->impl;
    for (int i = 0; i < fc::kNP; ++i)
        if (!std::isfinite(params[i])) return FC_ERR_INVALID;
    const std::vector<float> zeros(fc::kNP, 0.0f);
    const std::size_t pb = fc::kNP * sizeof(float);
    if (hipMemcpyAsync(m.dP, params, pb, hipMemcpyHostToDevice, m.stream) != hipSuccess ||
        hipMemcpyAsync(m.dM, zeros.data(), pb, hipMemcpyHostToDevice, m.stream) != hipSuccess ||
        hipMemcpyAsync(m.dV, zeros.data(), pb, hipMemcpyHostToDevice, m.stream) != hipSuccess ||
        hipStreamSynchronize(m.stream) != hipSuccess)
        return FC_ERR_DEVICE;
    m.step = 0;
    return FC_OK;
}

int fc_mlp_trainer_data(fc_mlp_trainer* t, float* x, float* target) {
    if (!t) return FC_ERR_INVALID;
    fc::MlpTrainer& m = t->impl;
    if ((x && hipMemcpyAsync(x, m.dX, 3 * m.n * sizeof(float), hipMemcpyDeviceToHost, m.stream) != hipSuccess) ||
        (target && hipMemcpyAsync(target, m.dT, m.n * sizeof(float), hipMemcpyDeviceToHost, m.stream) != hipSuccess) ||
        hipStreamSynchronize(m.stream) != hipSuccess)
        return FC_ERR_DEVICE;
    return FC_OK;
}

int fc_mlp_trainer_export(fc_mlp_trainer* t, fc_mlp_params* out) {
    if (!t || !out) return FC_ERR_INVALID;
    fc::MlpTrainer& m = t->impl;
    std::vector<float> P(fc::kNP);
    const int rc = fc_mlp_trainer_get_params(t, P.data());
    if (rc != FC_OK) return rc;
    std::memset(out, 0, sizeof(*out));
    for (int c = 0; c < 3; ++c) {
        out->in_mean[c] = m.inMean[c];
        out->in_inv_std[c] = m.inInvStd[c];
    }
    for (int c = 0; c < 4; ++c) {
        out->env_lo[c] = m.envLo[c];
        out->env_hi[c] = m.envHi[c];
    }
    out->out_mean = m.outMean;
    out->out_std = m.outStd;
    for (int i = 0; i < 96; ++i) out->w1[i] = fc::bf16BitsHost(P[fc::kOW1 + i]);
    for (int i = 0; i < 32; ++i) out->b1[i] = P[fc::kOB1 + i];
    for (int i = 0; i < 1024; ++i) out->w2[i] = fc::bf16BitsHost(P[fc::kOW2 + i]);
    for (int i = 0; i < 32; ++i) out->b2[i] = P[fc::kOB2 + i];
    for (int i = 0; i < 32; ++i) out->w3[i] = fc::bf16BitsHost(P[fc::kOW3 + i]);
    out->b3 = P[fc::kOB3];
    return FC_OK;
}

int fc_mlp_infer(const fc_mlp_params* params, const float* samples, int count, float* out) {
    if (!params || count < 0 || (count > 0 && (!samples || !out))) return FC_ERR_INVALID;
    if (count == 0) return FC_OK;
    fc::MlpParams* dParams = nullptr;
    float *dS = nullptr, *dO = nullptr;
    hipStream_t st = nullptr;
    int rc = FC_OK;
    if (!fc::devAlloc(&dParams, 1) || !fc::devAlloc(&dS, 4 * static_cast<std::size_t>(count)) ||
        !fc::devAlloc(&dO, static_cast<std::size_t>(count)) || hipStreamCreate(&st) != hipSuccess) {
        rc = FC_ERR_ALLOC;
    } else if (hipMemcpyAsync(dParams, params, sizeof(fc::MlpParams), hipMemcpyHostToDevice, st) != hipSuccess ||
               hipMemcpyAsync(dS, samples, 16 * static_cast<std::size_t>(count), hipMemcpyHostToDevice, st) !=
                   hipSuccess ||
               fc::launchMlpInfer(dParams, dS, count, dO, st) != 0 ||
               hipMemcpyAsync(out, dO, 4 * static_cast<std::size_t>(count), hipMemcpyDeviceToHost, st) !=
                   hipSuccess ||
               hipStreamSynchronize(st) != hipSuccess) {
        rc = FC_ERR_DEVICE;
    }
    if (st) {
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
    }
    if (rc != FC_OK) (void)hipGetLastError();
    fc::devFree(dParams);
    fc::devFree(dS);
    fc::devFree(dO);
    return rc;
}

}  // extern "C"
