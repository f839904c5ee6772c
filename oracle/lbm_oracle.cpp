// FoilCFD CPU oracle -- TEST INFRASTRUCTURE ONLY.
//
// A from-scratch CPU restatement of the reference solver's per-step
// arithmetic (Kemerd/foil_cfd src/sim/lbm_core.cu), written from the
// mathematics and the reference sources read as text.  It is the checker for
// the HIP kernels: only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it.  The product library never links it.
//
// Semantics follow the reference exactly:
//   * pull scheme, ping-pong buffers, post-collision populations stored;
//   * canonical layout f[q * ncells + cell], cell = x + nx*(y + ny*z); the
//     reference's two z ghost planes are replaced by an explicit z wrap, which
//     is what the ghost refresh after every step computes
//     (lbm_core.cu:484-497, lbm_solver.cpp:1164-1166);
//   * optional periodic x / y (not in the reference, which has no periodic
//     box mode; used for the Taylor-Green case).
// Everything is templated on the arithmetic type T (float = FP32 oracle,
// double = FP64 oracle).  Built with -ffp-contract=off so the FP32 path keeps
// the reference's operation order without fused multiply-adds.
//
// Parity status per function (what pins it) is stated at each function.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>
#include <algorithm>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int kQ = 19;
// D3Q19 frozen ordering, lbm_core.cuh:34-68.
constexpr int kCx[kQ] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0};
constexpr int kCy[kQ] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1};
constexpr int kCz[kQ] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1};
constexpr int kOpp[kQ] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
constexpr int kMirY[kQ] = {0, 1, 2, 4, 3, 5, 6, 13, 14, 9, 10, 18, 17, 7, 8, 15, 16, 12, 11};
constexpr int kMirZ[kQ] = {0, 1, 2, 3, 4, 6, 5, 7, 8, 15, 16, 17, 18, 13, 14, 9, 10, 11, 12};

template <class T> T weight(int q) {
    if (q == 0) return T(1) / T(3);
    if (q <= 6) return T(1) / T(18);
    return T(1) / T(36);
}

// CellFlag values, lbm_core.cuh:114-130.
constexpr uint8_t kFluid = 0, kSolid = 1, kInlet = 2, kOutlet = 3;
constexpr uint8_t kSlipT = 4, kSlipB = 5, kSlipF = 6, kSlipK = 7;

struct Grid {
    int nx, ny, nz;
    int px, py;  // periodic x / y
    long long ncells() const { return (long long)nx * ny * nz; }
    // Neighbour coordinate along an axis; returns -1 when it leaves a
    // non-periodic axis (z always wraps).
    static int step(int v, int d, int n, bool periodic) {
        int r = v + d;
        if (r < 0) return periodic ? r + n : -1;
        if (r >= n) return periodic ? r - n : -1;
        return r;
    }
    long long idx(int x, int y, int z) const {
        return (long long)x + (long long)nx * ((long long)y + (long long)ny * z);
    }
};

// IEEE binary16 -> binary32, exact (the reference decodes the slip field with
// __half2float, lbm_core.cu:374-376, which is exact).
float halfBitsToFloat(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    uint32_t exp = (h >> 10) & 0x1Fu;
    uint32_t man = h & 0x3FFu;
    uint32_t bits;
    if (exp == 0) {
        if (man == 0) {
            bits = sign;
        } else {  // subnormal: renormalise
            int e = -1;
            do { ++e; man <<= 1; } while ((man & 0x400u) == 0);
            man &= 0x3FFu;
            bits = sign | ((uint32_t)(127 - 15 - e) << 23) | (man << 13);
        }
    } else if (exp == 31) {
        bits = sign | 0x7F800000u | (man << 13);
    } else {
        bits = sign | ((exp + 127 - 15) << 23) | (man << 13);
    }
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}

// equilibrium19, lbm_core.cu:65-73.
template <class T>
void equilibrium19(T rho, T ux, T uy, T uz, T feq[kQ]) {
    const T usq = T(1.5) * (ux * ux + uy * uy + uz * uz);
    for (int q = 0; q < kQ; ++q) {
        const T cu = T(3) * (T(kCx[q]) * ux + T(kCy[q]) * uy + T(kCz[q]) * uz);
        feq[q] = weight<T>(q) * rho * (T(1) + cu + T(0.5) * cu * cu - usq);
    }
}

// initEquilibriumKernel, lbm_core.cu:102-118 (real cells only; the ghost
// planes it also writes are the z wrap here).
template <class T>
void initEquilibrium(T* f, const uint8_t* flags, const Grid& g, T uInlet) {
    const long long n = g.ncells();
#pragma omp parallel for schedule(static)
    for (long long c = 0; c < n; ++c) {
        const bool solid = flags[c] == kSolid;
        T feq[kQ];
        equilibrium19<T>(T(1), solid ? T(0) : uInlet, T(0), T(0), feq);
        for (int q = 0; q < kQ; ++q) f[(long long)q * n + c] = feq[q];
    }
}

// spanwisePerturbKernel + phase hash, lbm_core.cu:130-160, :763-767.
// The reference uses the __sinf intrinsic (F9); this uses the libm sine, so
// the field agrees to the intrinsic's error only (parity unpinned at that
// last-bits boundary; the integer phase hash is exact).
uint32_t speckPhaseBits(uint32_t seed, uint32_t salt) {
    uint32_t h = (seed + salt) * 2654435761u;
    h ^= h >> 16;
    return h & 0xFFFFFFu;
}

template <class T>
void spanwisePerturb(T* f, const uint8_t* flags, const Grid& g, float amplitude,
                     uint32_t seed) {
    const float p1 = 6.2831853f * (float)speckPhaseBits(seed, 1u) / 16777216.0f;
    const float p2 = 6.2831853f * (float)speckPhaseBits(seed, 2u) / 16777216.0f;
    const float p3 = 6.2831853f * (float)speckPhaseBits(seed, 3u) / 16777216.0f;
    const long long n = g.ncells();
#pragma omp parallel for schedule(static)
    for (int z = 0; z < g.nz; ++z) {
        for (int y = 0; y < g.ny; ++y) {
            for (int x = 0; x < g.nx; ++x) {
                const long long c = g.idx(x, y, z);
                if (flags[c] != kFluid) continue;
                const float tz = 6.2831853f * ((float)z + 0.5f) / (float)g.nz;
                const float tx = 6.2831853f * ((float)x + 0.5f) / (float)g.nx;
                const float wz = amplitude
                               * (0.6f * std::sin(3.0f * tz + p1)
                                  + 0.4f * std::sin(7.0f * tz + p2))
                               * (1.0f + 0.3f * std::sin(2.0f * tx + p3));
                for (int q = 0; q < kQ; ++q) {
                    if (kCz[q] == 0) continue;
                    T& v = f[(long long)q * n + c];
                    v = v + T(3) * weight<T>(q) * T(kCz[q]) * T(wz);
                }
            }
        }
    }
}

struct StepArgs {
    double tau0, magicLambda, smagPrefactor, uInlet;
    int writeMacro;
};

// streamCollideKernel<WM, QLIBB>, lbm_core.cu:213-475 (all four
// instantiations: the slip and q-link views are optional).
template <class T, class M>
void streamCollide(const T* src, T* dst, const uint8_t* flags, const Grid& g,
                   const StepArgs& a, M* mRho, M* mU, M* mV, M* mW,
                   const uint16_t* uwx, const uint16_t* uwy, const uint16_t* uwz,
                   const uint8_t* qFrac, const uint32_t* ffMask) {
    const long long n = g.ncells();
    const bool wm = uwx && uwy && uwz;
    const bool ql = qFrac && ffMask;
    const bool macro = a.writeMacro && mRho && mU && mV && mW;
    const T tau0 = T(a.tau0), lam = T(a.magicLambda), pre = T(a.smagPrefactor);
    const T uIn = T(a.uInlet);
    const T maxU = T(0.30f);  // kMaxSimSpeed, lbm_core.cuh:80

#pragma omp parallel for schedule(static)
    for (int z = 0; z < g.nz; ++z) {
        for (int y = 0; y < g.ny; ++y) {
            for (int x = 0; x < g.nx; ++x) {
                const long long cell = g.idx(x, y, z);
                const uint8_t my = flags[cell];
                if (my != kFluid) {
                    if (my == kInlet) {  // lbm_core.cu:242-253
                        T feq[kQ];
                        equilibrium19<T>(T(1), uIn, T(0), T(0), feq);
                        for (int q = 0; q < kQ; ++q) dst[(long long)q * n + cell] = feq[q];
                        if (macro) {
                            mRho[cell] = M(1); mU[cell] = M(uIn); mV[cell] = M(0); mW[cell] = M(0);
                        }
                    } else if (my == kOutlet) {  // lbm_core.cu:254-272
                        const int xm = Grid::step(x, -1, g.nx, g.px);
                        const long long up = g.idx(xm, y, z);
                        T r = 0, jx = 0, jy = 0, jz = 0;
                        for (int q = 0; q < kQ; ++q) {
                            const T fq = src[(long long)q * n + up];
                            dst[(long long)q * n + cell] = fq;
                            r += fq;
                            jx += T(kCx[q]) * fq; jy += T(kCy[q]) * fq; jz += T(kCz[q]) * fq;
                        }
                        if (macro) {
                            const T inv = T(1) / r;
                            mRho[cell] = M(r); mU[cell] = M(jx * inv);
                            mV[cell] = M(jy * inv); mW[cell] = M(jz * inv);
                        }
                    } else if (macro) {  // lbm_core.cu:273-282
                        mRho[cell] = M(1); mU[cell] = M(0); mV[cell] = M(0); mW[cell] = M(0);
                    }
                    continue;
                }

                // Fluid hot path, lbm_core.cu:303-383.
                T uwX = 0, uwY = 0, uwZ = 0;
                bool uwLoaded = false;
                T fpop[kQ];
                for (int q = 0; q < kQ; ++q) {
                    const int xn = Grid::step(x, -kCx[q], g.nx, g.px);
                    const int yn = Grid::step(y, -kCy[q], g.ny, g.py);
                    const int zn = Grid::step(z, -kCz[q], g.nz, true);
                    // Fluid cells never sit on a non-periodic face (the face
                    // planes are boundary-flagged); xn/yn are in range.
                    const long long pn = g.idx(xn, yn, zn);
                    const uint8_t nf = flags[pn];
                    const bool solid = nf == kSolid;
                    const bool slipY = nf == kSlipT || nf == kSlipB;
                    const bool slipZ = nf == kSlipF || nf == kSlipK;
                    int qs = q;
                    long long cs = pn;
                    if (solid) { qs = kOpp[q]; cs = cell; }
                    else if (slipY) { qs = kMirY[q]; cs = g.idx(xn, y, zn); }
                    else if (slipZ) { qs = kMirZ[q]; cs = g.idx(xn, yn, z); }
                    fpop[q] = src[(long long)qs * n + cs];

                    if (ql && solid) {  // lbm_core.cu:338-364 (as written, F6)
                        const uint8_t qb = qFrac[cell * kQ + q];
                        if (qb != 0) {
                            const T qf = T((float)qb * (1.0f / 255.0f));
                            const T fOut = src[(long long)q * n + cell];
                            if (qf >= T(0.5)) {
                                const T inv2q = T(0.5) / qf;
                                fpop[q] = inv2q * fOut + (T(1) - inv2q) * fpop[q];
                            } else if (ffMask[cell] & (1u << q)) {
                                const int xf = Grid::step(x, kCx[q], g.nx, g.px);
                                const int yf = Grid::step(y, kCy[q], g.ny, g.py);
                                const int zf = Grid::step(z, kCz[q], g.nz, true);
                                const T fOutFF = src[(long long)q * n + g.idx(xf, yf, zf)];
                                fpop[q] = T(2) * qf * fOut + (T(1) - T(2) * qf) * fOutFF;
                            }
                        }
                    }
                    if (wm && solid) {  // lbm_core.cu:366-382
                        if (!uwLoaded) {
                            uwX = T(halfBitsToFloat(uwx[cell]));
                            uwY = T(halfBitsToFloat(uwy[cell]));
                            uwZ = T(halfBitsToFloat(uwz[cell]));
                            uwLoaded = true;
                        }
                        fpop[q] += T(6) * weight<T>(q)
                                 * (T(kCx[q]) * uwX + T(kCy[q]) * uwY + T(kCz[q]) * uwZ);
                    }
                }

                // Moments, density floor, velocity limiter, lbm_core.cu:386-415.
                T rho = 0, jx = 0, jy = 0, jz = 0;
                for (int q = 0; q < kQ; ++q) {
                    rho += fpop[q];
                    jx += T(kCx[q]) * fpop[q];
                    jy += T(kCy[q]) * fpop[q];
                    jz += T(kCz[q]) * fpop[q];
                }
                rho = std::max(rho, T(0.05f));
                if (std::isnan(rho)) rho = T(0.05f);  // fmaxf(NaN, 0.05) = 0.05
                const T invRho = T(1) / rho;
                T ux = jx * invRho, uy = jy * invRho, uz = jz * invRho;
                const T u2 = ux * ux + uy * uy + uz * uz;
                if (u2 > maxU * maxU) {
                    const T s = maxU / std::sqrt(u2);
                    ux *= s; uy *= s; uz *= s;
                }

                T feq[kQ];
                equilibrium19<T>(rho, ux, uy, uz, feq);

                // Smagorinsky tau_eff, lbm_core.cu:427-444.
                T pxx = 0, pyy = 0, pzz = 0, pxy = 0, pxz = 0, pyz = 0;
                for (int q = 0; q < kQ; ++q) {
                    const T fneq = fpop[q] - feq[q];
                    pxx += T(kCx[q] * kCx[q]) * fneq;
                    pyy += T(kCy[q] * kCy[q]) * fneq;
                    pzz += T(kCz[q] * kCz[q]) * fneq;
                    pxy += T(kCx[q] * kCy[q]) * fneq;
                    pxz += T(kCx[q] * kCz[q]) * fneq;
                    pyz += T(kCy[q] * kCz[q]) * fneq;
                }
                T tauEff = tau0;
                if (pre > T(0)) {
                    const T qbar = std::sqrt(pxx * pxx + pyy * pyy + pzz * pzz
                                             + T(2) * (pxy * pxy + pxz * pxz + pyz * pyz));
                    tauEff = T(0.5) * (tau0 + std::sqrt(tau0 * tau0 + pre * qbar * invRho));
                }
                // TRT, lbm_core.cu:450-467.
                const T omP = T(1) / tauEff;
                const T tauM = T(0.5) + lam / (tauEff - T(0.5));
                const T omM = T(1) / tauM;
                dst[cell] = fpop[0] - omP * (fpop[0] - feq[0]);
                for (int q = 1; q < kQ; q += 2) {
                    const int b = q + 1;
                    const T sym = omP * (T(0.5) * (fpop[q] + fpop[b]) - T(0.5) * (feq[q] + feq[b]));
                    const T asym = omM * (T(0.5) * (fpop[q] - fpop[b]) - T(0.5) * (feq[q] - feq[b]));
                    dst[(long long)q * n + cell] = fpop[q] - sym - asym;
                    dst[(long long)b * n + cell] = fpop[b] - sym + asym;
                }
                if (macro) {
                    mRho[cell] = M(rho); mU[cell] = M(ux); mV[cell] = M(uy); mW[cell] = M(uz);
                }
            }
        }
    }
}

// forceReductionKernel<WM>, lbm_core.cu:573-637, with the per-link terms in
// T and the global sum in FP64 in a fixed (cell-ascending) order.
template <class T>
void forceReduce(const T* f, const uint8_t* flags, const Grid& g,
                 const uint16_t* uwx, const uint16_t* uwy, const uint16_t* uwz,
                 double out[3]) {
    const long long n = g.ncells();
    const bool wm = uwx && uwy && uwz;
    std::vector<double> part(3 * (size_t)g.nz, 0.0);
#pragma omp parallel for schedule(static)
    for (int z = 0; z < g.nz; ++z) {
        double sx = 0, sy = 0, sz = 0;
        for (int y = 0; y < g.ny; ++y) {
            for (int x = 0; x < g.nx; ++x) {
                const long long cell = g.idx(x, y, z);
                if (flags[cell] != kFluid) continue;
                T uwX = 0, uwY = 0, uwZ = 0;
                if (wm) {
                    uwX = T(halfBitsToFloat(uwx[cell]));
                    uwY = T(halfBitsToFloat(uwy[cell]));
                    uwZ = T(halfBitsToFloat(uwz[cell]));
                }
                for (int q = 1; q < kQ; ++q) {
                    const int xn = Grid::step(x, kCx[q], g.nx, g.px);
                    const int yn = Grid::step(y, kCy[q], g.ny, g.py);
                    const int zn = Grid::step(z, kCz[q], g.nz, true);
                    if (flags[g.idx(xn, yn, zn)] != kSolid) continue;
                    const T fq = f[(long long)q * n + cell];
                    T corr = 0;
                    if (wm)
                        corr = T(6) * weight<T>(q)
                             * (T(kCx[q]) * uwX + T(kCy[q]) * uwY + T(kCz[q]) * uwZ);
                    const T m = T(2) * fq - corr;
                    sx += (double)(T(kCx[q]) * m);
                    sy += (double)(T(kCy[q]) * m);
                    sz += (double)(T(kCz[q]) * m);
                }
            }
        }
        part[3 * (size_t)z + 0] = sx;
        part[3 * (size_t)z + 1] = sy;
        part[3 * (size_t)z + 2] = sz;
    }
    out[0] = out[1] = out[2] = 0.0;
    for (int z = 0; z < g.nz; ++z) {
        out[0] += part[3 * (size_t)z + 0];
        out[1] += part[3 * (size_t)z + 1];
        out[2] += part[3 * (size_t)z + 2];
    }
}

// nanWatchdogKernel, lbm_core.cu:522-559, :856-871: verdict of one sampled
// sweep (0 healthy, base+1 non-finite, base+2 velocity runaway).
template <class T>
int watchdog(const T* f, const Grid& g, long long sampleOffset, int flagBase) {
    const long long stride = 4093;
    const long long n = g.ncells();
    const long long off = ((sampleOffset % stride) + stride) % stride;
    int verdict = 0;
    for (long long cell = off; cell < n; cell += stride) {
        float s = 0, jx = 0, jy = 0, jz = 0;
        for (int q = 0; q < kQ; ++q) {
            const float fq = (float)f[(long long)q * n + cell];
            s += fq; jx += (float)kCx[q] * fq; jy += (float)kCy[q] * fq; jz += (float)kCz[q] * fq;
        }
        if (!std::isfinite(s)) {
            verdict = flagBase + 1;
        } else {
            const float speed2 = (jx * jx + jy * jy + jz * jz) / std::max(s * s, 1e-12f);
            const float cap = 0.95f * 0.30f;
            if (speed2 > cap * cap) verdict = flagBase + 2;
        }
    }
    return verdict;
}

template <class T>
void momentsOf(const T* f, const Grid& g, double* rho, double* u, double* v, double* w) {
    const long long n = g.ncells();
#pragma omp parallel for schedule(static)
    for (long long c = 0; c < n; ++c) {
        double r = 0, jx = 0, jy = 0, jz = 0;
        for (int q = 0; q < kQ; ++q) {
            const double fq = (double)f[(long long)q * n + c];
            r += fq; jx += kCx[q] * fq; jy += kCy[q] * fq; jz += kCz[q] * fq;
        }
        rho[c] = r; u[c] = jx / r; v[c] = jy / r; w[c] = jz / r;
    }
}


// ---- iMEM wall model (lbm_wallmodel.cu:31-196) -----------------------------
// Reichardt law, Newton inversion, EMA, linear slip balance, fade, clamp, FP16
// encoding. The reference uses the __logf / __expf intrinsics; this uses libm
// (parity to the intrinsics' error, not to the last bit).
constexpr float kKappa = 0.41f, kReichC = 7.8f, kFadeY0 = 3.0f, kFadeY1 = 6.0f;

float reichardtUPlus(float yp) {
    return std::log(1.0f + kKappa * yp) / kKappa
         + kReichC * (1.0f - std::exp(-yp / 11.0f) - (yp / 11.0f) * std::exp(-yp / 3.0f));
}

float reichardtDuPlus(float yp) {
    return 1.0f / (1.0f + kKappa * yp)
         + kReichC * ((1.0f / 11.0f) * std::exp(-yp / 11.0f) - (1.0f / 11.0f) * std::exp(-yp / 3.0f)
                      + (yp / 33.0f) * std::exp(-yp / 3.0f));
}

float solveUTauIters(float ut, float ys, float nu, float seed, int iters) {
    float utau = (seed > 0.0f) ? seed : std::max(0.04f * ut, std::sqrt(nu * ut / ys));
    for (int it = 0; it < iters; ++it) {
        const float yp = std::max(ys * utau / nu, 1e-8f);
        const float up = reichardtUPlus(yp);
        const float g = utau * up - ut;
        const float dg = up + yp * reichardtDuPlus(yp);
        utau -= g / std::max(dg, 1e-6f);
        utau = std::max(utau, 0.0f);
    }
    return utau;
}

float solveUTau(float ut, float ys, float nu, float seed) { return solveUTauIters(ut, ys, nu, seed, 6); }

// IEEE binary32 -> binary16, round to nearest even (= __float2half_rn).
uint16_t floatToHalfBits(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t absx = x & 0x7FFFFFFFu;
    if (absx >= 0x7F800000u)  // inf / nan
        return static_cast<uint16_t>(sign | 0x7C00u | (absx > 0x7F800000u ? 0x200u : 0u));
    if (absx >= 0x477FF000u) return static_cast<uint16_t>(sign | 0x7C00u);  // rounds to inf
    if (absx < 0x38800000u) {  // subnormal half (or zero)
        if (absx < 0x33000000u) return static_cast<uint16_t>(sign);  // < 2^-25 -> 0
        const uint32_t e = absx >> 23;
        const uint32_t m = (absx & 0x7FFFFFu) | 0x800000u;
        const uint32_t shift = 126u - e;  // 14 + (113 - e)
        const uint32_t q = m >> shift;
        const uint32_t rem = m & ((1u << shift) - 1u);
        const uint32_t half = 1u << (shift - 1);
        uint32_t r = q + ((rem > half || (rem == half && (q & 1u))) ? 1u : 0u);
        return static_cast<uint16_t>(sign | r);
    }
    uint32_t r = absx - 0x38000000u;  // rebias exponent (127 -> 15)
    const uint32_t q = r >> 13;
    const uint32_t rem = r & 0x1FFFu;
    uint32_t h = q + ((rem > 0x1000u || (rem == 0x1000u && (q & 1u))) ? 1u : 0u);
    return static_cast<uint16_t>(sign | h);
}

template <class T>
void wallUpdate(const T* f, const Grid& g, int count, const long long* cellIdx,
                const long long* sampleIdx, const float* nxA, const float* nyA, const float* nzA,
                const float* ysA, const uint32_t* maskA, float* uTau, float nu, float cutoff,
                float ema, uint16_t* uwx, uint16_t* uwy, uint16_t* uwz, float* stats4) {
    const long long n = g.ncells();
    float sumY = 0, maxY = 0;
    int active = 0, clamped = 0;
    for (int i = 0; i < count; ++i) {
        const long long cell = cellIdx[i];
        const float nX = nxA[i], nY = nyA[i], nZ = nzA[i];
        const long long sc = sampleIdx[i];
        float rho = 0, jx = 0, jy = 0, jz = 0;
        for (int q = 0; q < kQ; ++q) {
            const float fq = (float)f[(long long)q * n + sc];
            rho += fq; jx += kCx[q] * fq; jy += kCy[q] * fq; jz += kCz[q] * fq;
        }
        rho = std::max(rho, 0.05f);
        const float inv = 1.0f / rho;
        const float ux = jx * inv, uy = jy * inv, uz = jz * inv;
        const float un = ux * nX + uy * nY + uz * nZ;
        const float tx = ux - un * nX, ty = uy - un * nY, tz = uz - un * nZ;
        const float ut = std::sqrt(tx * tx + ty * ty + tz * tz);
        if (ut < cutoff) {
            uTau[i] = 0.0f;
            uwx[cell] = uwy[cell] = uwz[cell] = 0;
            continue;
        }
        const float itx = tx / ut, ity = ty / ut, itz = tz / ut;
        const float prev = uTau[i];
        const float solved = solveUTau(ut, ysA[i], nu, prev);
        const float utau = (prev > 0.0f) ? prev + ema * (solved - prev) : solved;
        uTau[i] = utau;
        float m0 = 0, dc = 0;
        for (int q = 1; q < kQ; ++q) {
            if (!(maskA[i] & (1u << q))) continue;
            const int o = kOpp[q];
            const float ct = -(kCx[q] * itx + kCy[q] * ity + kCz[q] * itz);
            m0 += 2.0f * (float)f[(long long)o * n + cell] * ct;
            dc += 6.0f * weight<float>(q) * ct * ct;
        }
        const float maxN = std::max(std::fabs(nX), std::max(std::fabs(nY), std::fabs(nZ)));
        const float aw = 1.0f / std::max(maxN, 0.1f);
        float uw = 0.0f;
        if (dc > 1e-6f) {
            const float target = rho * utau * utau * aw;
            const float sgn = (m0 >= 0.0f) ? 1.0f : -1.0f;
            uw = (m0 - sgn * target) / dc;
        }
        const float yps = ysA[i] * utau / nu;
        const float fade = std::min(std::max((yps - kFadeY0) / (kFadeY1 - kFadeY0), 0.0f), 1.0f);
        uw *= fade;
        const float cap = std::min(0.5f * ut, 0.5f * 0.30f);
        const float uwc = std::min(std::max(uw, -cap), cap);
        uwx[cell] = floatToHalfBits(uwc * itx);
        uwy[cell] = floatToHalfBits(uwc * ity);
        uwz[cell] = floatToHalfBits(uwc * itz);
        sumY += yps;
        maxY = std::max(maxY, yps);
        ++active;
        if (uwc != uw) ++clamped;
    }
    if (stats4) {
        stats4[0] = sumY; stats4[1] = maxY; stats4[2] = (float)active; stats4[3] = (float)clamped;
    }
}


// ---- tracer particles (particles.cu:48-376) ---------------------------------
// lowbias32 hash, 24-bit uniform draws, trilinear velocity at p - 0.5 (x / y
// clamped, z periodic), RK2 midpoint, respawn rules. Function arguments of
// the reference's make_float3/4 calls are taken as evaluated left to right.
uint32_t hashU32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352du;
    x ^= x >> 15; x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
float rand01(uint32_t& s) {
    s = hashU32(s);
    return (float)(s & 0x00FFFFFFu) * (1.0f / 16777216.0f);
}
int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

struct P3 { float x, y, z; };
P3 lerp3(P3 a, P3 b, float t) {
    return P3{a.x + (b.x - a.x) * t, a.y + (b.y - a.y) * t, a.z + (b.z - a.z) * t};
}
struct VelField {
    const float *u, *v, *w;
    int nx, ny, nz;
    P3 at(int x, int y, int z) const {
        int zz = z < 0 ? z + nz : (z >= nz ? z - nz : z);
        const long long c = (long long)clampi(x, 0, nx - 1) + (long long)nx * ((long long)clampi(y, 0, ny - 1) + (long long)ny * zz);
        return P3{u[c], v[c], w[c]};
    }
    P3 sample(P3 p) const {
        const float gx = std::min(std::max(p.x - 0.5f, 0.0f), (float)nx - 1.0001f);
        const float gy = std::min(std::max(p.y - 0.5f, 0.0f), (float)ny - 1.0001f);
        float gz = p.z - 0.5f;
        if (gz < 0.0f) gz += (float)nz;
        const int x0 = (int)gx, y0 = (int)gy, z0 = (int)gz;
        const float fx = gx - (float)x0, fy = gy - (float)y0, fz = gz - (float)z0;
        const P3 x00 = lerp3(at(x0, y0, z0), at(x0 + 1, y0, z0), fx);
        const P3 x10 = lerp3(at(x0, y0 + 1, z0), at(x0 + 1, y0 + 1, z0), fx);
        const P3 x01 = lerp3(at(x0, y0, z0 + 1), at(x0 + 1, y0, z0 + 1), fx);
        const P3 x11 = lerp3(at(x0, y0 + 1, z0 + 1), at(x0 + 1, y0 + 1, z0 + 1), fx);
        return lerp3(lerp3(x00, x10, fy), lerp3(x01, x11, fy), fz);
    }
};
bool finite3(P3 a) { return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z); }
P3 inletSpawn(int ny, int nz, uint32_t& r) {
    const float a = 0.5f + 1.5f * rand01(r);
    const float b = (float)ny * rand01(r);
    const float c = (float)nz * rand01(r);
    return P3{a, b, c};
}

Grid makeGrid(int nx, int ny, int nz, int px, int py) { return Grid{nx, ny, nz, px, py}; }

// ---------------------------------------------------------------------------
// Two-level grid refinement (lbm_refine.cu / lbm_refine.cuh, stepN sequencing
// lbm_solver.cpp:1172-1290). The fine level covers the coarse patch box
// [x0,x1) x [y0,y1) x [0,nz) at factor m (fine dims m*W, m*H, m*nz), x / y not
// periodic, z periodic; its x / y faces carry a 2-cell Interface shell.
// ---------------------------------------------------------------------------
constexpr uint8_t kIface = 8;
constexpr int kRestrictBand = 2;   // lbm_refine.cuh:77
constexpr int kRestrictBlend = 3;  // lbm_refine.cuh:85

struct PBox {
    int x0, y0, x1, y1;
};

// Latt-Chopard projection of fneq onto the second-order Hermite space,
// lbm_refine.cu:64-95. Pinned by: idempotence, zero mass / momentum of the
// result, and exact reproduction of a pure second-order fneq (tests).
template <class T>
void regularizeFneq(T fneq[kQ]) {
    T pxx = 0, pyy = 0, pzz = 0, pxy = 0, pxz = 0, pyz = 0;
    for (int q = 0; q < kQ; ++q) {
        const T f = fneq[q];
        const T cx = T(kCx[q]), cy = T(kCy[q]), cz = T(kCz[q]);
        pxx += cx * cx * f; pyy += cy * cy * f; pzz += cz * cz * f;
        pxy += cx * cy * f; pxz += cx * cz * f; pyz += cy * cz * f;
    }
    const T trace = T(1.0 / 3.0) * (pxx + pyy + pzz);
    for (int q = 0; q < kQ; ++q) {
        const T cx = T(kCx[q]), cy = T(kCy[q]), cz = T(kCz[q]);
        const T qPi = cx * cx * pxx + cy * cy * pyy + cz * cz * pzz +
                      T(2) * (cx * cy * pxy + cx * cz * pxz + cy * cz * pyz) - trace;
        fneq[q] = T(4.5) * weight<T>(q) * qPi;
    }
}

// Moments with the collision's density floor and speed cap
// (lbm_refine.cu:211-226 / :300-311).
template <class T>
void cappedMoments(const T fi[kQ], T& rho, T& ux, T& uy, T& uz) {
    T r = 0, jx = 0, jy = 0, jz = 0;
    for (int q = 0; q < kQ; ++q) {
        r += fi[q];
        jx += T(kCx[q]) * fi[q];
        jy += T(kCy[q]) * fi[q];
        jz += T(kCz[q]) * fi[q];
    }
    r = std::max(r, T(0.05f));
    const T inv = T(1) / r;
    ux = jx * inv; uy = jy * inv; uz = jz * inv;
    const T u2 = ux * ux + uy * uy + uz * uz;
    const T cap = T(0.30f);
    if (u2 > cap * cap) {
        const T sc = cap / std::sqrt(u2);
        ux *= sc; uy *= sc; uz *= sc;
    }
    rho = r;
}

// coarseToFineFillKernel, lbm_refine.cu:123-258: fine Interface cells (or,
// fullVolume, every non-solid fine cell) get the coarse post-collision
// populations trilinearly interpolated at their centre (solid corners dropped
// and weights renormalised), blended between coarse t (cT0) and t+1 (cT1),
// re-split into feq + regularised fneq, fneq rescaled by neqScale.
//
// Test switch (orc_set_shell_poison): when set, the shell fill stores NaN in
// every population f_q of a shell cell x whose target x + c_q lies outside the
// shell's interior (x, y in [2, n - 2)). A run that stays finite and bit-
// identical with it set shows that the next step, a nested level's band and
// the restriction read only the populations that stream into the interior.
static int g_shellPoison = 0;

template <class T>
void coarseToFineFill(const T* cT0, const T* cT1, const uint8_t* cFlags, const Grid& cg,
                      const uint8_t* fFlags, T* ff, const Grid& fg, const PBox& box, int m,
                      T timeWeight, T neqScale, bool fullVolume) {
    const long long fn = fg.ncells(), cn = cg.ncells();
    const T invM = T(1) / T(m);
#pragma omp parallel for schedule(static)
    for (long long cell = 0; cell < fn; ++cell) {
        const uint8_t flag = fFlags[cell];
        if (fullVolume ? flag == kSolid : flag != kIface) continue;
        const int xf = (int)(cell % fg.nx);
        const int yf = (int)((cell / fg.nx) % fg.ny);
        const int zf = (int)(cell / ((long long)fg.nx * fg.ny));
        const T gx = T(box.x0) + (T(xf) + T(0.5)) * invM - T(0.5);
        const T gy = T(box.y0) + (T(yf) + T(0.5)) * invM - T(0.5);
        const T gz = (T(zf) + T(0.5)) * invM - T(0.5);
        const int ix = (int)std::floor(gx), iy = (int)std::floor(gy), iz = (int)std::floor(gz);
        const T tx = gx - T(ix), ty = gy - T(iy), tz = gz - T(iz);
        T wgt[8] = {(1 - tx) * (1 - ty) * (1 - tz), tx * (1 - ty) * (1 - tz),
                    (1 - tx) * ty * (1 - tz),       tx * ty * (1 - tz),
                    (1 - tx) * (1 - ty) * tz,       tx * (1 - ty) * tz,
                    (1 - tx) * ty * tz,             tx * ty * tz};
        long long corner[8];
        for (int c = 0; c < 8; ++c) {
            const int cx = std::min(std::max(ix + (c & 1), 0), cg.nx - 1);
            const int cy = std::min(std::max(iy + ((c >> 1) & 1), 0), cg.ny - 1);
            int cz = iz + ((c >> 2) & 1);
            if (cz < 0) cz += cg.nz;
            if (cz >= cg.nz) cz -= cg.nz;
            corner[c] = cg.idx(cx, cy, cz);
        }
        T wsum = 0;
        for (int c = 0; c < 8; ++c) {
            if (cFlags[corner[c]] == kSolid) wgt[c] = 0;
            wsum += wgt[c];
        }
        if (wsum < T(1e-6f)) continue;
        const T invW = T(1) / wsum;
        for (int c = 0; c < 8; ++c) wgt[c] *= invW;
        T fi[kQ];
        for (int q = 0; q < kQ; ++q) {
            T s0 = 0, s1 = 0;
            for (int c = 0; c < 8; ++c) {
                s0 += wgt[c] * cT0[(long long)q * cn + corner[c]];
                s1 += wgt[c] * cT1[(long long)q * cn + corner[c]];
            }
            fi[q] = s0 + timeWeight * (s1 - s0);
        }
        T rho, ux, uy, uz;
        cappedMoments(fi, rho, ux, uy, uz);
        T feq[kQ], fneq[kQ];
        equilibrium19<T>(rho, ux, uy, uz, feq);
        for (int q = 0; q < kQ; ++q) fneq[q] = fi[q] - feq[q];
        regularizeFneq(fneq);
        for (int q = 0; q < kQ; ++q) {
            T v = feq[q] + neqScale * fneq[q];
            if (!fullVolume && g_shellPoison) {
                const int tx = xf + kCx[q], ty = yf + kCy[q];
                if (tx < 2 || ty < 2 || tx >= fg.nx - 2 || ty >= fg.ny - 2) v = std::numeric_limits<T>::quiet_NaN();
            }
            ff[(long long)q * fn + cell] = v;
        }
    }
}

// childAvgFneq, lbm_refine.cu:274-317: mean of the m^3 fine children
// (solid children skipped), capped moments, fneq = mean - feq.
template <class T>
bool childAvgFneq(const T* ff, const uint8_t* fFlags, const Grid& fg, int fx, int fy, int fz, int m,
                  T fneqOut[kQ], T& rho, T& ux, T& uy, T& uz) {
    const long long fn = fg.ncells();
    T fi[kQ];
    for (int q = 0; q < kQ; ++q) fi[q] = 0;
    int nf = 0;
    for (int k = 0; k < m; ++k)
        for (int j = 0; j < m; ++j)
            for (int i = 0; i < m; ++i) {
                const long long c = fg.idx(fx + i, fy + j, fz + k);
                if (fFlags[c] == kSolid) continue;
                ++nf;
                for (int q = 0; q < kQ; ++q) fi[q] += ff[(long long)q * fn + c];
            }
    if (nf == 0) return false;
    const T invN = T(1) / T(nf);
    for (int q = 0; q < kQ; ++q) fi[q] *= invN;
    cappedMoments(fi, rho, ux, uy, uz);
    T feq[kQ];
    equilibrium19<T>(rho, ux, uy, uz, feq);
    for (int q = 0; q < kQ; ++q) fneqOut[q] = fi[q] - feq[q];
    return true;
}

// fineToCoarseRestrictKernel, lbm_refine.cu:319-429: coarse Fluid cells of the
// box shrunk by kRestrictBand get the child-averaged state with the 19-point
// fneq filter (neighbours inside the restricted x / y range with fluid
// children; z periodic), regularised, rescaled by neqScaleInv, blended into
// the existing coarse post-collision value over a kRestrictBlend-cell ramp.
template <class T, class M>
void fineToCoarseRestrict(const T* ff, const uint8_t* fFlags, const Grid& fg, T* fc,
                          const uint8_t* cFlags, const Grid& cg, const PBox& box, int m,
                          T neqScaleInv, M* mRho, M* mU, M* mV, M* mW) {
    const int rx0 = box.x0 + kRestrictBand, ry0 = box.y0 + kRestrictBand;
    const int rx1 = box.x1 - kRestrictBand, ry1 = box.y1 - kRestrictBand;
    const long long cn = cg.ncells();
    const bool macro = mRho && mU && mV && mW;
#pragma omp parallel for schedule(static) collapse(2)
    for (int zc = 0; zc < cg.nz; ++zc)
        for (int yc = ry0; yc < ry1; ++yc)
            for (int xc = rx0; xc < rx1; ++xc) {
                const long long cc = cg.idx(xc, yc, zc);
                if (cFlags[cc] != kFluid) continue;
                const int dEdge = std::min(std::min(xc - rx0, rx1 - 1 - xc),
                                           std::min(yc - ry0, ry1 - 1 - yc));
                const T blend = std::min(T(1), T((float)(dEdge + 1) / (float)(kRestrictBlend + 1)));
                T fneqC[kQ], rho, ux, uy, uz;
                if (!childAvgFneq(ff, fFlags, fg, m * (xc - box.x0), m * (yc - box.y0), m * zc, m,
                                  fneqC, rho, ux, uy, uz))
                    continue;
                T fneqF[kQ];
                for (int q = 0; q < kQ; ++q) fneqF[q] = fneqC[q];
                int nS = 1;
                for (int nq = 1; nq < kQ; ++nq) {
                    const int nxc = xc + kCx[nq], nyc = yc + kCy[nq];
                    if (nxc < rx0 || nxc >= rx1 || nyc < ry0 || nyc >= ry1) continue;
                    const int nzw = (zc + kCz[nq] + cg.nz) % cg.nz;
                    T fneqN[kQ], rn, un, vn, wn;
                    if (!childAvgFneq(ff, fFlags, fg, m * (nxc - box.x0), m * (nyc - box.y0), m * nzw,
                                      m, fneqN, rn, un, vn, wn))
                        continue;
                    for (int q = 0; q < kQ; ++q) fneqF[q] += fneqN[q];
                    ++nS;
                }
                const T invS = T(1) / T(nS);
                for (int q = 0; q < kQ; ++q) fneqF[q] *= invS;
                regularizeFneq(fneqF);
                T feq[kQ];
                equilibrium19<T>(rho, ux, uy, uz, feq);
                T br = 0, bjx = 0, bjy = 0, bjz = 0;
                for (int q = 0; q < kQ; ++q) {
                    T& slot = fc[(long long)q * cn + cc];
                    const T restricted = feq[q] + neqScaleInv * fneqF[q];
                    const T blended = slot + blend * (restricted - slot);
                    slot = blended;
                    br += blended;
                    bjx += T(kCx[q]) * blended;
                    bjy += T(kCy[q]) * blended;
                    bjz += T(kCz[q]) * blended;
                }
                if (macro) {
                    br = std::max(br, T(0.05f));
                    const T inv = T(1) / br;
                    mRho[cc] = M(br); mU[cc] = M(bjx * inv); mV[cc] = M(bjy * inv); mW[cc] = M(bjz * inv);
                }
            }
}

// fineTauFor, lbm_refine.cuh:96-98 (FP32 as the reference computes it).
float fineTauFor(float tauC, int m) { return (float)m * (tauC - 0.5f) + 0.5f; }

// Optional third level (the nested "finer" patch): grid, flags, buffers and
// box in fine cells.
template <class T>
struct FinerLevel {
    T* gA = nullptr;
    T* gB = nullptr;
    const uint8_t* flags = nullptr;
    Grid grid{};
    PBox box{};
    int m2 = 0;
};

// The refined step sequence of stepN (lbm_solver.cpp:1144-1290): coarse step
// t -> t+1; for each of the m fine sub-steps the Interface fill from coarse
// (t, t+1) at weight k/m, then the fine step; with a finer level, m2 finer
// sub-steps follow each fine step (fill from fine t / t+1 at q/m2, finer
// step) and the restriction of the finer state into the fine t+1 buffer,
// before the fine swap (lbm_solver.cpp:1209-1267); then the restriction of
// the fine state into the coarse t+1 buffer. Taus and rescale factors are
// formed in FP32 as the reference does.
template <class T>
void runRefined(T* cA, T* cB, const uint8_t* cFlags, const Grid& cg, T* fA, T* fB,
                const uint8_t* fFlags, const Grid& fg, const PBox& box, int m, int nsteps,
                const T* taus, const T* uins, T lam, T pre, T* rho, T* u, T* v, T* w,
                const FinerLevel<T>* finer = nullptr, const uint16_t* const* cSlip = nullptr,
                const uint16_t* const* fSlip = nullptr, const uint8_t* const* qFracs = nullptr,
                const uint32_t* const* ffMasks = nullptr) {
    // optional q-LIBB fields per level (coarse, fine), as the reference's
    // streamCollideKernel<WM, QLIBB> takes each level's QLinkView
    const uint8_t* cq = qFracs ? qFracs[0] : nullptr;
    const uint8_t* fq = qFracs ? qFracs[1] : nullptr;
    const uint32_t* cff = ffMasks ? ffMasks[0] : nullptr;
    const uint32_t* fff = ffMasks ? ffMasks[1] : nullptr;
    // optional wall-model slip fields (half bits) of the coarse and fine
    // levels, held fixed over the run as within one stepN batch
    const uint16_t* cux = cSlip ? cSlip[0] : nullptr;
    const uint16_t* cuy = cSlip ? cSlip[1] : nullptr;
    const uint16_t* cuz = cSlip ? cSlip[2] : nullptr;
    const uint16_t* fux = fSlip ? fSlip[0] : nullptr;
    const uint16_t* fuy = fSlip ? fSlip[1] : nullptr;
    const uint16_t* fuz = fSlip ? fSlip[2] : nullptr;
    T* cs = cA;
    T* cd = cB;
    T* fs = fA;
    T* fd = fB;
    T* gs = finer ? finer->gA : nullptr;
    T* gd = finer ? finer->gB : nullptr;
    for (int i = 0; i < nsteps; ++i) {
        const float tauC = (float)taus[i];
        const float tauF = fineTauFor(tauC, m);
        const bool last = i == nsteps - 1;
        StepArgs a{taus[i], lam, pre, uins[i], last};
        streamCollide<T, T>(cs, cd, cFlags, cg, a, rho, u, v, w, cux, cuy, cuz, cq, cff);
        StepArgs af{(T)tauF, lam, pre, uins[i], 0};
        const float neqScale = tauF / ((float)m * tauC);
        for (int k = 0; k < m; ++k) {
            coarseToFineFill<T>(cs, cd, cFlags, cg, fFlags, fs, fg, box, m,
                                (T)((float)k / (float)m), (T)neqScale, false);
            streamCollide<T, T>(fs, fd, fFlags, fg, af, (T*)nullptr, (T*)nullptr, (T*)nullptr,
                                (T*)nullptr, fux, fuy, fuz, fq, fff);
            if (finer) {
                const int m2 = finer->m2;
                const float tauG = fineTauFor(tauF, m2);
                StepArgs ag{(T)tauG, lam, pre, uins[i], 0};
                const float neqG = tauG / ((float)m2 * tauF);
                for (int q = 0; q < m2; ++q) {
                    coarseToFineFill<T>(fs, fd, fFlags, fg, finer->flags, gs, finer->grid, finer->box, m2,
                                        (T)((float)q / (float)m2), (T)neqG, false);
                    streamCollide<T, T>(gs, gd, finer->flags, finer->grid, ag, (T*)nullptr, (T*)nullptr,
                                        (T*)nullptr, (T*)nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
                    std::swap(gs, gd);
                }
                const float neqGInv = ((float)m2 * tauF) / tauG;
                fineToCoarseRestrict<T, T>(gs, finer->flags, finer->grid, fd, fFlags, fg, finer->box, m2,
                                           (T)neqGInv, (T*)nullptr, (T*)nullptr, (T*)nullptr, (T*)nullptr);
            }
            std::swap(fs, fd);
        }
        const float neqScaleInv = ((float)m * tauC) / tauF;
        fineToCoarseRestrict<T, T>(fs, fFlags, fg, cd, cFlags, cg, box, m, (T)neqScaleInv,
                                   last ? rho : nullptr, last ? u : nullptr, last ? v : nullptr,
                                   last ? w : nullptr);
        std::swap(cs, cd);
    }
    if (cs != cA) std::memcpy(cA, cs, sizeof(T) * (size_t)kQ * (size_t)cg.ncells());
    if (fs != fA) std::memcpy(fA, fs, sizeof(T) * (size_t)kQ * (size_t)fg.ncells());
    if (finer && gs != finer->gA)
        std::memcpy(finer->gA, gs, sizeof(T) * (size_t)kQ * (size_t)finer->grid.ncells());
}


// ---------------------------------------------------------------------------
// Learned u_tau network (no reference counterpart; the reference solves
// Reichardt's law by Newton only, lbm_wallmodel.cu:55-73). A restatement of
// the network the HIP kernels train and evaluate, written from the
// mathematics: 3 -> 32 -> 32 -> 1 with tanh hidden units, mean squared error,
// reverse-mode gradient, Adam (Kingma & Ba 2015). Packed master parameters
// (1217 floats): W1[32][3] @0, b1[32] @96, W2[32][32] @128 (row = output
// unit), b2[32] @1152, w3[32] @1184, b3 @1216.
// With emulate = true the operands the GPU feeds to BF16 matrix-core tiles
// are rounded to BF16 here too (weights; first-layer activations into the
// hidden layer; second-layer deltas and first-layer activations into the
// weight-gradient product; deltas and weights into the input-delta product),
// everything else in FP64. With emulate = false the whole network is FP64,
// and the gradient is the exact derivative of the computed loss.
// Parity pins: gradient vs central finite differences (FP64, emulate =
// false); forward vs a closed form with one active path; the first Adam step
// equals lr * sign(g).
constexpr int kMlpH = 32;
constexpr int kMlpNP = 1217;
constexpr int kOW1 = 0, kOB1 = 96, kOW2 = 128, kOB2 = 1152, kOW3 = 1184, kOB3 = 1216;

float bf16RoundF(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even at bit 16
    u &= 0xFFFF0000u;
    float r;
    std::memcpy(&r, &u, 4);
    return r;
}

inline double bfe(double v, bool emulate) { return emulate ? (double)bf16RoundF((float)v) : v; }

struct MlpTape {
    double h1[kMlpH], h2[kMlpH], y;
};

void mlpForward(const float* P, bool emu, const double x[3], MlpTape& t) {
    for (int k = 0; k < kMlpH; ++k) {
        double pre = P[kOB1 + k];
        for (int c = 0; c < 3; ++c) pre += bfe(P[kOW1 + 3 * k + c], emu) * x[c];
        t.h1[k] = std::tanh(pre);
    }
    for (int j = 0; j < kMlpH; ++j) {
        double z = P[kOB2 + j];
        for (int k = 0; k < kMlpH; ++k) z += bfe(P[kOW2 + j * kMlpH + k], emu) * bfe(t.h1[k], emu);
        t.h2[j] = std::tanh(z);
    }
    double y = P[kOB3];
    for (int j = 0; j < kMlpH; ++j) y += bfe(P[kOW3 + j], emu) * t.h2[j];
    t.y = y;
}

// grad = d/dP of (1/batch) sum_i (y_i - T_i)^2 over the samples idx[0..batch);
// returns the sum of squared errors. X is SoA [3][n].
double mlpGradient(const float* P, bool emu, const float* X, long long n, const float* T, const int* idx,
                   int batch, double* grad) {
    std::fill(grad, grad + kMlpNP, 0.0);
    double sse = 0.0;
    for (int b = 0; b < batch; ++b) {
        const long long g = idx[b];
        const double x[3] = {X[g], X[n + g], X[2 * n + g]};
        MlpTape t;
        mlpForward(P, emu, x, t);
        const double e = t.y - T[g];
        sse += e * e;
        const double dy = 2.0 * e / batch;
        grad[kOB3] += dy;
        double dz2[kMlpH];
        for (int j = 0; j < kMlpH; ++j) {
            grad[kOW3 + j] += dy * t.h2[j];
            dz2[j] = dy * bfe(P[kOW3 + j], emu) * (1.0 - t.h2[j] * t.h2[j]);
            grad[kOB2 + j] += dz2[j];
        }
        for (int j = 0; j < kMlpH; ++j)
            for (int k = 0; k < kMlpH; ++k) grad[kOW2 + j * kMlpH + k] += bfe(dz2[j], emu) * bfe(t.h1[k], emu);
        for (int k = 0; k < kMlpH; ++k) {
            double dh = 0.0;
            for (int j = 0; j < kMlpH; ++j) dh += bfe(dz2[j], emu) * bfe(P[kOW2 + j * kMlpH + k], emu);
            const double dz1 = dh * (1.0 - t.h1[k] * t.h1[k]);
            grad[kOB1 + k] += dz1;
            for (int c = 0; c < 3; ++c) grad[kOW1 + 3 * k + c] += dz1 * x[c];
        }
    }
    return sse;
}

void mlpAdam(float* P, float* M, float* V, const double* grad, double lr, double b1, double b2, double eps,
             int step) {
    const double c1 = 1.0 - std::pow(b1, step), c2 = 1.0 - std::pow(b2, step);
    for (int p = 0; p < kMlpNP; ++p) {
        const double m = b1 * M[p] + (1.0 - b1) * grad[p];
        const double v = b2 * V[p] + (1.0 - b2) * grad[p] * grad[p];
        M[p] = (float)m;
        V[p] = (float)v;
        P[p] = (float)(P[p] - lr * (m / c1) / (std::sqrt(v / c2) + eps));
    }
}

}  // namespace

extern "C" {

int orc_num_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

float orc_half_to_float(uint16_t h) { return halfBitsToFloat(h); }

void orc_equilibrium_f64(double rho, double ux, double uy, double uz, double* feq) {
    equilibrium19<double>(rho, ux, uy, uz, feq);
}
void orc_equilibrium_f32(float rho, float ux, float uy, float uz, float* feq) {
    equilibrium19<float>(rho, ux, uy, uz, feq);
}

void orc_init_equilibrium_f32(float* f, const uint8_t* flags, int nx, int ny, int nz,
                              float uInlet) {
    initEquilibrium<float>(f, flags, makeGrid(nx, ny, nz, 0, 0), uInlet);
}
void orc_init_equilibrium_f64(double* f, const uint8_t* flags, int nx, int ny, int nz,
                              double uInlet) {
    initEquilibrium<double>(f, flags, makeGrid(nx, ny, nz, 0, 0), uInlet);
}

void orc_spanwise_perturb_f32(float* f, const uint8_t* flags, int nx, int ny, int nz,
                              float amplitude, uint32_t seed) {
    spanwisePerturb<float>(f, flags, makeGrid(nx, ny, nz, 0, 0), amplitude, seed);
}
void orc_spanwise_perturb_f64(double* f, const uint8_t* flags, int nx, int ny, int nz,
                              float amplitude, uint32_t seed) {
    spanwisePerturb<double>(f, flags, makeGrid(nx, ny, nz, 0, 0), amplitude, seed);
}

uint32_t orc_speck_phase_bits(uint32_t seed, uint32_t salt) { return speckPhaseBits(seed, salt); }

// Smagorinsky prefactor exactly as the launch wrapper forms it in FP32,
// lbm_core.cu:785-787.
float orc_smag_prefactor(float cs) {
    return (cs > 0.0f) ? 18.0f * 1.41421356f * cs * cs : 0.0f;
}

void orc_stream_collide_f32(const float* src, float* dst, const uint8_t* flags,
                            int nx, int ny, int nz, int periodicX, int periodicY,
                            float tau0, float magicLambda, float smagPrefactor,
                            float uInlet, int writeMacro,
                            float* rho, float* u, float* v, float* w,
                            const uint16_t* uwx, const uint16_t* uwy, const uint16_t* uwz,
                            const uint8_t* qFrac, const uint32_t* ffMask) {
    StepArgs a{tau0, magicLambda, smagPrefactor, uInlet, writeMacro};
    streamCollide<float, float>(src, dst, flags, makeGrid(nx, ny, nz, periodicX, periodicY), a,
                                rho, u, v, w, uwx, uwy, uwz, qFrac, ffMask);
}

void orc_stream_collide_f64(const double* src, double* dst, const uint8_t* flags,
                            int nx, int ny, int nz, int periodicX, int periodicY,
                            double tau0, double magicLambda, double smagPrefactor,
                            double uInlet, int writeMacro,
                            double* rho, double* u, double* v, double* w,
                            const uint16_t* uwx, const uint16_t* uwy, const uint16_t* uwz,
                            const uint8_t* qFrac, const uint32_t* ffMask) {
    StepArgs a{tau0, magicLambda, smagPrefactor, uInlet, writeMacro};
    streamCollide<double, double>(src, dst, flags, makeGrid(nx, ny, nz, periodicX, periodicY), a,
                                  rho, u, v, w, uwx, uwy, uwz, qFrac, ffMask);
}

// n steps of the plain step sequence (stream-collide + ping-pong swap),
// with per-step tau and inlet speed supplied by the caller (the solver's
// startup ramps, lbm_solver.cpp:1144-1154). The macro arrays are written on
// the last step (writeMacro = i == n-1). fA holds the input state and, on
// return, the output state; fB holds the second ping-pong buffer and, on
// return, the state before the last step (the swap count is handled
// internally).
void orc_run_f32(float* fA, float* fB, const uint8_t* flags, int nx, int ny, int nz,
                 int periodicX, int periodicY, int nsteps, const float* tauPerStep,
                 const float* uInPerStep, float magicLambda, float smagPrefactor,
                 float* rho, float* u, float* v, float* w,
                 const uint16_t* uwx, const uint16_t* uwy, const uint16_t* uwz,
                 const uint8_t* qFrac, const uint32_t* ffMask) {
    const Grid g = makeGrid(nx, ny, nz, periodicX, periodicY);
    float* src = fA;
    float* dst = fB;
    for (int i = 0; i < nsteps; ++i) {
        StepArgs a{tauPerStep[i], magicLambda, smagPrefactor, uInPerStep[i], i == nsteps - 1};
        streamCollide<float, float>(src, dst, flags, g, a, rho, u, v, w, uwx, uwy, uwz, qFrac,
                                    ffMask);
        std::swap(src, dst);
    }
    if (src != fA) std::swap_ranges(fA, fA + (size_t)kQ * (size_t)g.ncells(), fB);
}

void orc_run_f64(double* fA, double* fB, const uint8_t* flags, int nx, int ny, int nz,
                 int periodicX, int periodicY, int nsteps, const double* tauPerStep,
                 const double* uInPerStep, double magicLambda, double smagPrefactor,
                 double* rho, double* u, double* v, double* w,
                 const uint16_t* uwx, const uint16_t* uwy, const uint16_t* uwz,
                 const uint8_t* qFrac, const uint32_t* ffMask) {
    const Grid g = makeGrid(nx, ny, nz, periodicX, periodicY);
    double* src = fA;
    double* dst = fB;
    for (int i = 0; i < nsteps; ++i) {
        StepArgs a{tauPerStep[i], magicLambda, smagPrefactor, uInPerStep[i], i == nsteps - 1};
        streamCollide<double, double>(src, dst, flags, g, a, rho, u, v, w, uwx, uwy, uwz,
                                      qFrac, ffMask);
        std::swap(src, dst);
    }
    if (src != fA) std::swap_ranges(fA, fA + (size_t)kQ * (size_t)g.ncells(), fB);
}

// Storage-format emulation (test infrastructure, no reference counterpart):
// the FP64 step sequence of orc_run_f64 with every population rounded after
// each step (and once on entry) as a device storage format keeps it: the
// deviation d = f - w_q scaled by `scale`, rounded to nearest-even in a
// binary format with `mant` stored mantissa bits and normals from 2^emin
// (subnormal quantum below), then unscaled. Formats: FP32 (23, -126, 1),
// IEEE FP16 with K = 2^15 (10, -14, 32768), BF16 (7, -126, 1).
static double roundFormat(double x, int mant, int emin) {
    if (x == 0.0 || !std::isfinite(x)) return x;
    int e = 0;
    std::frexp(x, &e);
    int ex = e - 1;
    if (ex < emin) ex = emin;
    const double q = std::ldexp(1.0, ex - mant);
    return std::nearbyint(x / q) * q;
}

// FP16C (s1e4m11) exactly as the device encodes it: the FP32 deviation is
// rescaled by 2^-112 and its bits rounded to nearest even at bit 12 (15-bit
// magnitude, saturating), then decoded back (restatement of the device's
// StoreF16C, which has no reference counterpart).
static uint16_t f16cEncode(float v) {
    // saturate at the largest code (1 + 2047/2048), NaN to the largest
    // negative code, then round to nearest even at bit 12 of the FP32 bits
    const float lim = 0x1.FFEp0f * 0x1p-112f;
    const float y = std::fmin(std::fmax(v * 0x1p-112f, -lim), lim);
    uint32_t bits;
    std::memcpy(&bits, &y, 4);
    const uint32_t r = bits + 0x7FFu + ((bits >> 12) & 1u);
    return static_cast<uint16_t>(((r >> 12) & 0x7FFFu) | ((r >> 16) & 0x8000u));
}
static float f16cDecode(uint16_t b) {
    const uint32_t x = b;
    const uint32_t u = ((x & 0x7FFFu) << 12) | ((x & 0x8000u) << 16);
    float y;
    std::memcpy(&y, &u, 4);
    return y * 0x1p112f;
}

// mant < 0 selects FP16C.
static void roundStorage(double* f, long long n, int mant, int emin, double scale) {
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < n; ++i) {
        for (int q = 0; q < kQ; ++q) {
            const double w = weight<double>(q);
            double& v = f[(long long)q * n + i];
            if (mant < 0)
                v = w + (double)f16cDecode(f16cEncode((float)(v - w)));
            else
                v = w + roundFormat((v - w) * scale, mant, emin) / scale;
        }
    }
}

void orc_run_stored_f64(double* fA, double* fB, const uint8_t* flags, int nx, int ny, int nz,
                        int periodicX, int periodicY, int nsteps, const double* tauPerStep,
                        const double* uInPerStep, double magicLambda, double smagPrefactor,
                        double* rho, double* u, double* v, double* w, int mant, int emin,
                        double scale, int keepOther) {
    const Grid g = makeGrid(nx, ny, nz, periodicX, periodicY);
    const long long n = g.ncells();
    double* src = fA;
    double* dst = fB;
    roundStorage(src, n, mant, emin, scale);
    if (keepOther)
        roundStorage(dst, n, mant, emin, scale);
    else
        std::memcpy(dst, src, sizeof(double) * (size_t)kQ * (size_t)n);
    for (int i = 0; i < nsteps; ++i) {
        StepArgs a{tauPerStep[i], magicLambda, smagPrefactor, uInPerStep[i], i == nsteps - 1};
        streamCollide<double, double>(src, dst, flags, g, a, rho, u, v, w, nullptr, nullptr,
                                      nullptr, nullptr, nullptr);
        roundStorage(dst, n, mant, emin, scale);
        std::swap(src, dst);
    }
    if (src != fA) std::memcpy(fA, src, sizeof(double) * (size_t)kQ * (size_t)n);
}

// ---- refinement wrappers ----
void orc_regularize_fneq_f64(double* fneq) { regularizeFneq<double>(fneq); }

void orc_set_shell_poison(int on) { g_shellPoison = on ? 1 : 0; }

void orc_fine_fill_f64(const double* cT0, const double* cT1, const uint8_t* cFlags, int cnx,
                       int cny, int cnz, const uint8_t* fFlags, double* ff, int fnx, int fny,
                       int fnz, int x0, int y0, int x1, int y1, int m, double timeWeight,
                       double neqScale, int fullVolume) {
    coarseToFineFill<double>(cT0, cT1, cFlags, makeGrid(cnx, cny, cnz, 0, 0), fFlags, ff,
                             makeGrid(fnx, fny, fnz, 0, 0), PBox{x0, y0, x1, y1}, m, timeWeight,
                             neqScale, fullVolume != 0);
}
void orc_fine_fill_f32(const float* cT0, const float* cT1, const uint8_t* cFlags, int cnx,
                       int cny, int cnz, const uint8_t* fFlags, float* ff, int fnx, int fny,
                       int fnz, int x0, int y0, int x1, int y1, int m, float timeWeight,
                       float neqScale, int fullVolume) {
    coarseToFineFill<float>(cT0, cT1, cFlags, makeGrid(cnx, cny, cnz, 0, 0), fFlags, ff,
                            makeGrid(fnx, fny, fnz, 0, 0), PBox{x0, y0, x1, y1}, m, timeWeight,
                            neqScale, fullVolume != 0);
}
void orc_fine_restrict_f64(const double* ff, const uint8_t* fFlags, int fnx, int fny, int fnz,
                           double* fc, const uint8_t* cFlags, int cnx, int cny, int cnz, int x0,
                           int y0, int x1, int y1, int m, double neqScaleInv, double* rho,
                           double* u, double* v, double* w) {
    fineToCoarseRestrict<double, double>(ff, fFlags, makeGrid(fnx, fny, fnz, 0, 0), fc, cFlags,
                                         makeGrid(cnx, cny, cnz, 0, 0), PBox{x0, y0, x1, y1}, m,
                                         neqScaleInv, rho, u, v, w);
}
float orc_fine_tau_for(float tauC, int m) { return fineTauFor(tauC, m); }

void orc_run_refined_f64(double* cA, double* cB, const uint8_t* cFlags, int cnx, int cny,
                         int cnz, double* fA, double* fB, const uint8_t* fFlags, int x0, int y0,
                         int x1, int y1, int m, int nsteps, const double* taus,
                         const double* uins, double lam, double pre, double* rho, double* u,
                         double* v, double* w) {
    const Grid cg = makeGrid(cnx, cny, cnz, 0, 0);
    const Grid fg = makeGrid(m * (x1 - x0), m * (y1 - y0), m * cnz, 0, 0);
    runRefined<double>(cA, cB, cFlags, cg, fA, fB, fFlags, fg, PBox{x0, y0, x1, y1}, m, nsteps,
                       taus, uins, lam, pre, rho, u, v, w);
}
// Two levels with the coarse and fine wall-model slip fields (3 half-bit
// arrays each, NULL = no slip on that level) and q-LIBB fields (cut bytes
// and two-node masks per level, NULL = plain bounce-back on that level).
void orc_run_refined_slip_f64(double* cA, double* cB, const uint8_t* cFlags, int cnx, int cny, int cnz,
                              double* fA, double* fB, const uint8_t* fFlags, int x0, int y0, int x1, int y1,
                              int m, int nsteps, const double* taus, const double* uins, double lam, double pre,
                              double* rho, double* u, double* v, double* w, const uint16_t* cux,
                              const uint16_t* cuy, const uint16_t* cuz, const uint16_t* fux, const uint16_t* fuy,
                              const uint16_t* fuz, const uint8_t* cq, const uint32_t* cff, const uint8_t* fq,
                              const uint32_t* fff) {
    const Grid cg = makeGrid(cnx, cny, cnz, 0, 0);
    const Grid fg = makeGrid(m * (x1 - x0), m * (y1 - y0), m * cnz, 0, 0);
    const uint16_t* cs[3] = {cux, cuy, cuz};
    const uint16_t* fs[3] = {fux, fuy, fuz};
    const uint8_t* qs[2] = {cff ? cq : nullptr, fff ? fq : nullptr};
    const uint32_t* ms[2] = {cq ? cff : nullptr, fq ? fff : nullptr};
    runRefined<double>(cA, cB, cFlags, cg, fA, fB, fFlags, fg, PBox{x0, y0, x1, y1}, m, nsteps, taus, uins, lam,
                       pre, rho, u, v, w, nullptr, cux ? cs : nullptr, fux ? fs : nullptr, qs, ms);
}
void orc_run_refined_f32(float* cA, float* cB, const uint8_t* cFlags, int cnx, int cny, int cnz,
                         float* fA, float* fB, const uint8_t* fFlags, int x0, int y0, int x1,
                         int y1, int m, int nsteps, const float* taus, const float* uins,
                         float lam, float pre, float* rho, float* u, float* v, float* w) {
    const Grid cg = makeGrid(cnx, cny, cnz, 0, 0);
    const Grid fg = makeGrid(m * (x1 - x0), m * (y1 - y0), m * cnz, 0, 0);
    runRefined<float>(cA, cB, cFlags, cg, fA, fB, fFlags, fg, PBox{x0, y0, x1, y1}, m, nsteps,
                      taus, uins, lam, pre, rho, u, v, w);
}

// Three levels: coarse, fine (box in coarse cells, factor m) and finer (box
// gx0..gy1 in fine cells, factor m2); g* buffers hold the finer populations.
void orc_run_refined3_f64(double* cA, double* cB, const uint8_t* cFlags, int cnx, int cny, int cnz,
                          double* fA, double* fB, const uint8_t* fFlags, int x0, int y0, int x1, int y1,
                          int m, double* gA, double* gB, const uint8_t* gFlags, int gx0, int gy0, int gx1,
                          int gy1, int m2, int nsteps, const double* taus, const double* uins, double lam,
                          double pre, double* rho, double* u, double* v, double* w) {
    const Grid cg = makeGrid(cnx, cny, cnz, 0, 0);
    const Grid fg = makeGrid(m * (x1 - x0), m * (y1 - y0), m * cnz, 0, 0);
    FinerLevel<double> g;
    g.gA = gA;
    g.gB = gB;
    g.flags = gFlags;
    g.grid = makeGrid(m2 * (gx1 - gx0), m2 * (gy1 - gy0), m2 * fg.nz, 0, 0);
    g.box = PBox{gx0, gy0, gx1, gy1};
    g.m2 = m2;
    runRefined<double>(cA, cB, cFlags, cg, fA, fB, fFlags, fg, PBox{x0, y0, x1, y1}, m, nsteps, taus, uins,
                       lam, pre, rho, u, v, w, &g);
}

uint16_t orc_f16c_encode(float v) { return f16cEncode(v); }
float orc_f16c_decode(uint16_t b) { return f16cDecode(b); }

void orc_force_f32(const float* f, const uint8_t* flags, int nx, int ny, int nz,
                   int periodicX, int periodicY, const uint16_t* uwx, const uint16_t* uwy,
                   const uint16_t* uwz, double* out3) {
    forceReduce<float>(f, flags, makeGrid(nx, ny, nz, periodicX, periodicY), uwx, uwy, uwz, out3);
}
void orc_force_f64(const double* f, const uint8_t* flags, int nx, int ny, int nz,
                   int periodicX, int periodicY, const uint16_t* uwx, const uint16_t* uwy,
                   const uint16_t* uwz, double* out3) {
    forceReduce<double>(f, flags, makeGrid(nx, ny, nz, periodicX, periodicY), uwx, uwy, uwz, out3);
}

int orc_watchdog_f32(const float* f, int nx, int ny, int nz, long long sampleOffset,
                     int flagBase) {
    return watchdog<float>(f, makeGrid(nx, ny, nz, 0, 0), sampleOffset, flagBase);
}

void orc_moments_f32(const float* f, int nx, int ny, int nz, double* rho, double* u, double* v,
                     double* w) {
    momentsOf<float>(f, makeGrid(nx, ny, nz, 0, 0), rho, u, v, w);
}
void orc_moments_f64(const double* f, int nx, int ny, int nz, double* rho, double* u, double* v,
                     double* w) {
    momentsOf<double>(f, makeGrid(nx, ny, nz, 0, 0), rho, u, v, w);
}


float orc_reichardt_uplus(float yp) { return reichardtUPlus(yp); }
float orc_solve_utau(float ut, float ys, float nu, float seed) { return solveUTau(ut, ys, nu, seed); }
uint16_t orc_float_to_half(float f) { return floatToHalfBits(f); }

void orc_wall_update_f32(const float* f, int nx, int ny, int nz, int count, const long long* cellIdx,
                         const long long* sampleIdx, const float* nxA, const float* nyA,
                         const float* nzA, const float* ys, const uint32_t* mask, float* uTau,
                         float nu, float cutoff, float ema, uint16_t* uwx, uint16_t* uwy,
                         uint16_t* uwz, float* stats4) {
    wallUpdate<float>(f, makeGrid(nx, ny, nz, 0, 0), count, cellIdx, sampleIdx, nxA, nyA, nzA, ys,
                      mask, uTau, nu, cutoff, ema, uwx, uwy, uwz, stats4);
}


void orc_particles_seed(float* pos4, int count, int nx, int ny, int nz, uint32_t seed) {
    for (int i = 0; i < count; ++i) {
        uint32_t r = hashU32(seed ^ ((uint32_t)i * 0x9E3779B9u));
        const float a = (float)nx * rand01(r);
        const float b = (float)ny * rand01(r);
        const float c = (float)nz * rand01(r);
        const float e = rand01(r);
        pos4[4 * i] = a; pos4[4 * i + 1] = b; pos4[4 * i + 2] = c; pos4[4 * i + 3] = e;
    }
}

void orc_particles_advect(float* pos4, int count, const float* u, const float* v, const float* w,
                          const uint8_t* flags, int nx, int ny, int nz, float dt, float ageRate,
                          float inletFrac, uint32_t seed) {
    const VelField vf{u, v, w, nx, ny, nz};
    const float nzf = (float)nz;
    auto cellOf = [&](int x, int y, int z) {
        return (long long)x + (long long)nx * ((long long)y + (long long)ny * z);
    };
    for (int i = 0; i < count; ++i) {
        uint32_t r = hashU32(seed ^ ((uint32_t)i * 0x9E3779B9u));
        P3 p{pos4[4 * i], pos4[4 * i + 1], pos4[4 * i + 2]};
        float age = pos4[4 * i + 3];
        bool blew = !finite3(p);
        if (!blew && dt > 0.0f) {
            const P3 u1 = vf.sample(p);
            P3 pm{p.x + 0.5f * dt * u1.x, p.y + 0.5f * dt * u1.y, p.z + 0.5f * dt * u1.z};
            if (pm.z < 0.0f) pm.z += nzf; else if (pm.z >= nzf) pm.z -= nzf;
            const P3 u2 = vf.sample(pm);
            p = P3{p.x + dt * u2.x, p.y + dt * u2.y, p.z + dt * u2.z};
            blew = !finite3(p) || !finite3(u2);
            age -= ageRate * dt;
        }
        if (!blew) {
            if (p.z < 0.0f) p.z += nzf; else if (p.z >= nzf) p.z -= nzf;
        }
        const bool outflow = blew || p.x >= (float)nx - 1.0f;
        const bool oob = !outflow && (p.x < 0.0f || p.y < 0.0f || p.y >= (float)ny);
        const int cx = clampi((int)p.x, 0, nx - 1), cy = clampi((int)p.y, 0, ny - 1),
                  cz = clampi((int)p.z, 0, nz - 1);
        const bool solid = !outflow && !oob && flags && flags[cellOf(cx, cy, cz)] == kSolid;
        if (outflow || oob || solid || age <= 0.0f) {
            const bool toInlet = outflow || oob || rand01(r) < inletFrac;
            if (toInlet) {
                p = inletSpawn(ny, nz, r);
            } else {
                bool done = false;
                for (int attempt = 0; attempt < 4 && !done; ++attempt) {
                    const float a = (float)nx * rand01(r);
                    const float b = (float)ny * rand01(r);
                    const float c = (float)nz * rand01(r);
                    if (!flags) { p = P3{a, b, c}; done = true; break; }
                    const long long cc = cellOf(clampi((int)a, 0, nx - 1), clampi((int)b, 0, ny - 1),
                                                clampi((int)c, 0, nz - 1));
                    if (flags[cc] != kSolid) { p = P3{a, b, c}; done = true; }
                }
                if (!done) p = inletSpawn(ny, nz, r);
            }
            age = 0.6f + 0.4f * rand01(r);
        }
        pos4[4 * i] = p.x; pos4[4 * i + 1] = p.y; pos4[4 * i + 2] = p.z; pos4[4 * i + 3] = age;
    }
}

float orc_solve_utau_iters(float ut, float ys, float nu, float seed, int iters) {
    return solveUTauIters(ut, ys, nu, seed, iters);
}

float orc_bf16_round(float x) { return bf16RoundF(x); }

void orc_mlp_forward(const float* P, int emulate, const float* X, long long n, double* y) {
    for (long long i = 0; i < n; ++i) {
        const double x[3] = {X[i], X[n + i], X[2 * n + i]};
        MlpTape t;
        mlpForward(P, emulate != 0, x, t);
        y[i] = t.y;
    }
}

double orc_mlp_gradient(const float* P, int emulate, const float* X, long long n, const float* T, const int* idx,
                        int batch, double* grad) {
    return mlpGradient(P, emulate != 0, X, n, T, idx, batch, grad);
}

void orc_mlp_adam(float* P, float* M, float* V, const double* grad, double lr, double b1, double b2, double eps,
                  int step) {
    mlpAdam(P, M, V, grad, lr, b1, b2, eps, step);
}

}  // extern "C"
