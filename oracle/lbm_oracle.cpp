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

Grid makeGrid(int nx, int ny, int nz, int px, int py) { return Grid{nx, ny, nz, px, py}; }

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
// return, the output state (the swap count is handled internally).
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
    if (src != fA) std::memcpy(fA, src, sizeof(float) * (size_t)kQ * (size_t)g.ncells());
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
    if (src != fA) std::memcpy(fA, src, sizeof(double) * (size_t)kQ * (size_t)g.ncells());
}

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

}  // extern "C"
