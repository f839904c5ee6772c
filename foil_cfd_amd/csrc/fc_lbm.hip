// FoilCFD D3Q19 kernels for gfx950: fused TRT + Smagorinsky stream-collide
// on an in-place (AA) population store, equilibrium init, spanwise speck
// perturbation, canonical upload/download, FP64 momentum-exchange force,
// FP64 global mass and the NaN watchdog.
//
// Behaviour mirrors the reference kernels (src/sim/lbm_core.cu) cell for cell;
// the implementation is organised for CDNA4 instead:
//  * one population array updated in place (AA pattern) instead of a ping-pong
//    pair: half the footprint, and a step reads + writes each population once;
//  * one 64-lane wavefront per 64-cell x segment of a (y,z) row, so all row /
//    plane offsets (including the periodic z wrap, which replaces the
//    reference's ghost planes and their refresh kernel) are wave-uniform
//    scalars and only the x index lives in a vector register;
//  * workgroups remapped so each XCD (private L2) streams one contiguous slab;
//  * collision in pair (symmetric / antisymmetric) form on scaled deviations
//    H = K (f - w), ~250 VALU ops per cell, so FP16 storage stays HBM bound;
//  * the zero-gradient outlet, which needs the previous post-collision state
//    of the x-1 column (overwritten in place), reads a 2 x 19 x ny*nz feed
//    plane that the x-1 cells publish every step.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "fc_kcommon.hpp"
#include "fc_refine_math.hpp"

// No implicit contraction in this file: fused multiply-adds are written out
// (fmac), so every variant of the step rounds identically.
#pragma clang fp contract(off)

namespace fc {
namespace {

// Scaled equilibrium deviation K * (feq_q(rho, u) - w_q) for one direction,
// via the same constants the collision uses.
template <class C>
__device__ inline C heqOf(int q, C K, C drho, C rho, C ux, C uy, C uz, C usq) {
    const C w = wOf<C>(q);
    const C cu = C(kCx[q]) * ux + C(kCy[q]) * uy + C(kCz[q]) * uz;
    return K * w * fmac(rho, fmac(C(4.5) * cu, cu, fmac(C(3), cu, -usq)), drho);
}

// ---------------------------------------------------------------------------
// TRT + Smagorinsky collision of one cell on scaled deviations H (in place).
// Same operator as the reference (src/sim/lbm_core.cu:385-475): moments,
// density floor 0.05, |u| limiter 0.30, second-order equilibrium, tau_eff
// from |Pi_neq| (Hou et al.), TRT with tau- from the magic parameter. The
// order of operations differs (pair form, moments from pair sums, Pi_neq from
// the closed-form equilibrium moments), so FP32 results agree to rounding.
// ---------------------------------------------------------------------------
template <class C>
struct Macro {
    C rho, ux, uy, uz;
};

template <class C>
__device__ inline Macro<C> collide(C H[kQ], C K, C tau0, C lam, C pre) {
    const C invK = C(1) / K;
    C s[9], d[9];
#pragma unroll
    for (int p = 0; p < 9; ++p) {
        s[p] = H[2 * p + 1] + H[2 * p + 2];
        d[p] = H[2 * p + 1] - H[2 * p + 2];
    }
    // Pairs: 0 x, 1 y, 2 z, 3 (+x+y), 4 (+x+z), 5 (+y+z), 6 (+x-y), 7 (+x-z), 8 (+y-z).
    const C R = H[0] + ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7])) + s[8];
    const C jx = (d[0] + d[3]) + (d[4] + d[6]) + d[7];
    const C jy = (d[1] + d[3]) + (d[5] - d[6]) + d[8];
    const C jz = (d[2] + d[4]) + (d[5] - d[7]) - d[8];

    C drho = R * invK;
    C rho = C(1) + drho;
    if (!(rho >= C(0.05))) {  // fmax(rho, 0.05) incl. the NaN scrub
        rho = C(0.05);
        drho = C(-0.95);
    }
    const C invRho = frcp(rho);
    const C jscale = invK * invRho;
    C ux = jx * jscale, uy = jy * jscale, uz = jz * jscale;
    C u2 = fmac(ux, ux, fmac(uy, uy, uz * uz));
    if (u2 > C(0.09)) {  // kMaxSimSpeed^2
        const C sc = C(0.3) * frsqrt(u2);
        ux *= sc;
        uy *= sc;
        uz *= sc;
        u2 = fmac(ux, ux, fmac(uy, uy, uz * uz));
    }
    const C usq = C(1.5) * u2;

    // Smagorinsky effective relaxation time.
    C tauEff = tau0;
    if (pre > C(0)) {
        const C third = drho * C(1.0 / 3.0);
        const C rux = rho * ux, ruy = rho * uy, ruz = rho * uz;
        const C nxx = fmac(-rux, ux, fmac((s[0] + s[3]) + (s[4] + s[6]) + s[7], invK, -third));
        const C nyy = fmac(-ruy, uy, fmac((s[1] + s[3]) + (s[5] + s[6]) + s[8], invK, -third));
        const C nzz = fmac(-ruz, uz, fmac((s[2] + s[4]) + (s[5] + s[7]) + s[8], invK, -third));
        const C nxy = fmac(-rux, uy, (s[3] - s[6]) * invK);
        const C nxz = fmac(-rux, uz, (s[4] - s[7]) * invK);
        const C nyz = fmac(-ruy, uz, (s[5] - s[8]) * invK);
        const C off = fmac(nxy, nxy, fmac(nxz, nxz, nyz * nyz));
        const C q2 = fmac(nxx, nxx, fmac(nyy, nyy, fmac(nzz, nzz, C(2) * off)));
        const C qbar = fsqrt(q2);
        tauEff = C(0.5) * (tau0 + fsqrt(fmac(tau0, tau0, pre * qbar * invRho)));
    }
    const C omP = frcp(tauEff);
    const C omM = frcp(fmac(lam, frcp(tauEff - C(0.5)), C(0.5)));
    const C P = C(0.5) * (C(1) - omP);
    const C M = C(0.5) * (C(1) - omM);

    // Equilibrium constants per weight class (w = 1/18 axis, 1/36 diagonal):
    // sym part A + B cu^2, antisym part Cc cu, pre-multiplied by omega.
    const C base = fmac(-rho, usq, drho);
    const C a1 = omP * K * C(1.0 / 18.0) * base;
    const C b1 = omP * K * C(4.5 / 18.0) * rho;
    const C c1 = omM * K * C(3.0 / 18.0) * rho;
    const C a2 = C(0.5) * a1, b2 = C(0.5) * b1, c2 = C(0.5) * c1;
    const C a0 = omP * K * C(1.0 / 3.0) * base;

    H[0] = fmac(C(1) - omP, H[0], a0);
    const C cu[9] = {ux, uy, uz, ux + uy, ux + uz, uy + uz, ux - uy, ux - uz, uy - uz};
#pragma unroll
    for (int p = 0; p < 9; ++p) {
        const C A = p < 3 ? a1 : a2;
        const C B = p < 3 ? b1 : b2;
        const C Cc = p < 3 ? c1 : c2;
        const C c = cu[p];
        const C Sp = fmac(P, s[p], fmac(B, c * c, A));
        const C Dp = fmac(M, d[p], Cc * c);
        H[2 * p + 1] = Sp + Dp;
        H[2 * p + 2] = Sp - Dp;
    }
    return Macro<C>{rho, ux, uy, uz};
}

// An empty asm that takes v as input and output: the load that produced v
// has completed at this point, and later loads cannot be hoisted above it.
template <class R>
__device__ inline void pinLoaded(R& v) {
    asm volatile("" : "+v"(v));
}

// Slot of the incoming population q of this cell in the current layout: from
// L, f*_opp(q) of the upstream cell; from S, the cell's own slot q.
template <bool FROM_L, int q>
struct PullSlot {
    static constexpr int dir = FROM_L ? kOpp[q] : q;
    static constexpr int ex = FROM_L ? -kCx[q] : 0;
    static constexpr int ey = FROM_L ? -kCy[q] : 0;
    static constexpr int ez = FROM_L ? -kCz[q] : 0;
};

// ---------------------------------------------------------------------------
// The fused step. FROM_L: the array is in layout L and this step writes
// layout S (scatter writes); otherwise it reads layout S and writes L (fully
// local). WM compiles in the moving-wall (Ladd) slip term on solid links.
// M is the population accessor (PopBuf for arrays under 2 GiB, else PopPtr).
// ---------------------------------------------------------------------------
// KEEP: a fluid cell's H[q] is left holding the value as stored (rounded
// and decoded), for callers that use the stored state afterwards.
template <class S, template <class> class M, bool FROM_L, bool WM, bool PX, bool PY, bool KEEP = false>
__device__ inline void cellBody(const KArgs& a, const RowPos& p, const M<S>& pop, std::uint8_t k,
                                typename S::C (&H)[kQ]) {
    using T = typename S::T;
    using C = typename S::C;
    const unsigned c = nb(p, 0, 0, 0);
    const std::uint8_t flag = k & kKindFlagMask;
    const C K = C(S::K);
    const int newLayout = FROM_L ? kLayoutS : kLayoutL;
    const unsigned yz = static_cast<unsigned>(p.y + a.ny * p.z);
    const long long plane = static_cast<long long>(a.ny) * a.nz;

    // Store f*_q of this cell into the output layout (scatter writes are
    // bounds checked only for boundary cells; fluid cells never leave the
    // domain because every non-periodic face plane is boundary-flagged).
    // 16-bit storages round stochastically when the launch asks for it, with
    // the cell's random bits for this step.
    constexpr bool kNarrow = sizeof(T) == 2;
    const bool sto = kNarrow && a.stochastic;
    const unsigned rh = kNarrow ? roundHash(c, a.roundSeed) : 0u;
    // Returns the 16-bit formats' stored bits (unused otherwise).
    auto storeOut = [&](auto qc, C val, bool check) -> typename S::Bits {
        constexpr int q = decltype(qc)::value;
        if constexpr (kNarrow) {
            const typename S::Bits b = roundedBits<S>(val, sto, roundBits(rh, q));
            if (FROM_L) {
                if (check && !inDomain<PX>(a, p, kCx[q], kCy[q])) return b;
                pop.stBits(q, kCx[q], kCy[q], kCz[q], b);
            } else {
                pop.stBits(kOpp[q], 0, 0, 0, b);
            }
            return b;
        } else {
            if (FROM_L) {
                if (check && !inDomain<PX>(a, p, kCx[q], kCy[q])) return 0;
                pop.st(q, kCx[q], kCy[q], kCz[q], val);
            } else {
                pop.st(kOpp[q], 0, 0, 0, val);
            }
            return 0;
        }
    };

    if (flag != kFluid) {
        if (flag == kInlet) {  // equilibrium Dirichlet, reference lbm_core.cu:242-253
            const C uIn = C(a.uIn);
            const C usq = C(1.5) * uIn * uIn;
            staticFor<kQ>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                storeOut(qc, heqOf<C>(q, K, C(0), C(1), uIn, C(0), C(0), usq), true);
            });
            if (a.writeMacro) {
                a.rho[c] = 1.0f;
                a.u[c] = a.uIn;
                a.v[c] = 0.0f;
                a.w[c] = 0.0f;
            }
        } else if (flag == kOutlet) {  // zero-gradient copy, lbm_core.cu:254-272
            // H holds the feed plane's values (loaded with the pulls)
            C r = 0, jx = 0, jy = 0, jz = 0;
            staticFor<kQ>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                const C h = H[q];
                storeOut(qc, h, true);
                r += h;
                jx += C(kCx[q]) * h;
                jy += C(kCy[q]) * h;
                jz += C(kCz[q]) * h;
            });
            if (a.writeMacro) {
                const C rr = C(1) + r / K;
                const C inv = C(1) / rr;
                a.rho[c] = static_cast<float>(rr);
                a.u[c] = static_cast<float>(jx / K * inv);
                a.v[c] = static_cast<float>(jy / K * inv);
                a.w[c] = static_cast<float>(jz / K * inv);
            }
        } else if (a.writeMacro) {  // solid / slip / interface: no population write
            a.rho[c] = 1.0f;
            a.u[c] = 0.0f;
            a.v[c] = 0.0f;
            a.w[c] = 0.0f;
        }
        return;
    }

    // ---- fluid: H holds the 19 pulled populations --------------------------
    if (k & kKindSpecial) {
        // Per-link source select by the upstream neighbour's flag, as the
        // reference (lbm_core.cu:306-331), mapped onto the AA slots.
        C uwx = 0, uwy = 0, uwz = 0;
        if (WM && (k & kKindWall)) {
            uwx = C(halfBits(a.uwx[c]));
            uwy = C(halfBits(a.uwy[c]));
            uwz = C(halfBits(a.uwz[c]));
        }
        if constexpr (std::is_same_v<M<S>, PopPtr<S>>) {
            // The 18 upstream kinds in one batch, folded into three link
            // masks as they arrive; then the 18 source slots (bounce-back, a
            // mirror, or the plain pull again, picked without a branch) in
            // three groups of six, each group completed before the next (the
            // empty asm) so at most six selected addresses are live. A wave
            // holding a wall-adjacent cell waits for four memory round trips
            // instead of one or two per link.
            unsigned solidM = 0, slipYM = 0, slipZM = 0;
            staticFor<kQ - 1>([&](auto qm) {
                constexpr int q = decltype(qm)::value + 1;
                const std::uint8_t nf = a.kind[nb(p, -kCx[q], -kCy[q], -kCz[q])] & kKindFlagMask;
                solidM |= static_cast<unsigned>(nf == kSolid) << q;
                slipYM |= static_cast<unsigned>(nf == kSlipTop || nf == kSlipBottom) << q;
                slipZM |= static_cast<unsigned>(nf == kSlipFront || nf == kSlipBack) << q;
            });
            // NARROW: the slot's element index in 32 bits (kQ * ncells <
            // 2^32), else (direction slab, cell) with a 64-bit product. The
            // narrow form takes 73 instead of 83-90 VGPRs in the plain-wall
            // L->S step and is faster there; in the other variants the wide
            // form uses fewer registers (62-72) and is faster (measured on
            // MI355X), so they take it on every grid.
            auto fetchLinks = [&](auto narrowTag) {
                constexpr bool NARROW = decltype(narrowTag)::value;
                staticFor<3>([&](auto gc) {
                    constexpr int g = decltype(gc)::value;
                    staticFor<6>([&](auto jc) {
                        constexpr int q = 1 + 6 * g + decltype(jc)::value;
                        constexpr int cx = kCx[q], cy = kCy[q], cz = kCz[q];
                        constexpr int my = kMirY[q], mz = kMirZ[q];
                        using L = PullSlot<FROM_L, q>;
                        const bool solid = (solidM >> q) & 1u;
                        const bool slipY = (slipYM >> q) & 1u;
                        const bool slipZ = (slipZM >> q) & 1u;
                        // (direction slab, cell) of bounce-back, y mirror, z mirror, plain pull
                        constexpr int dBB = FROM_L ? q : kOpp[q], dSY = FROM_L ? kOpp[my] : my;
                        constexpr int dSZ = FROM_L ? kOpp[mz] : mz;
                        const unsigned cBB = FROM_L ? nb(p, 0, 0, 0) : nb(p, -cx, -cy, -cz);
                        const unsigned cSY = FROM_L ? nb(p, -cx, 0, -cz) : nb(p, 0, -cy, 0);
                        const unsigned cSZ = FROM_L ? nb(p, -cx, -cy, 0) : nb(p, 0, 0, -cz);
                        const unsigned cPL = nb(p, L::ex, L::ey, L::ez);
                        if constexpr (NARROW) {
                            const unsigned n32 = static_cast<unsigned>(pop.n);
                            const unsigned iBB = dBB * n32 + cBB, iSY = dSY * n32 + cSY;
                            const unsigned iSZ = dSZ * n32 + cSZ, iPL = L::dir * n32 + cPL;
                            H[q] = S::load(pop.f + (solid ? iBB : (slipY ? iSY : (slipZ ? iSZ : iPL))));
                        } else {
                            const int d = solid ? dBB : (slipY ? dSY : (slipZ ? dSZ : L::dir));
                            const unsigned cc = solid ? cBB : (slipY ? cSY : (slipZ ? cSZ : cPL));
                            H[q] = S::load(pop.f + static_cast<long long>(d) * pop.n + cc);
                        }
                    });
                    staticFor<6>([&](auto jc) { pinLoaded(H[1 + 6 * g + decltype(jc)::value]); });
                });
            };
            if constexpr (WM || !FROM_L)
                fetchLinks(std::false_type{});
            else if (!a.wideIndex)
                fetchLinks(std::true_type{});
            else
                fetchLinks(std::false_type{});
            if (WM) {
                staticFor<kQ - 1>([&](auto qm) {
                    constexpr int q = decltype(qm)::value + 1;
                    const C cuw = C(kCx[q]) * uwx + C(kCy[q]) * uwy + C(kCz[q]) * uwz;
                    const C withSlip = fmac(K * C(6) * wOf<C>(q), cuw, H[q]);
                    H[q] = ((solidM >> q) & 1u) ? withSlip : H[q];
                });
            }
        } else staticFor<kQ - 1>([&](auto qm) {
            constexpr int q = decltype(qm)::value + 1;
            constexpr int cx = kCx[q], cy = kCy[q], cz = kCz[q];
            const std::uint8_t nf = a.kind[nb(p, -cx, -cy, -cz)] & kKindFlagMask;
            if (nf == kSolid) {
                // half-way bounce-back: own f*_opp(q)
                if (FROM_L)
                    H[q] = pop.ld(q, 0, 0, 0);
                else
                    H[q] = pop.ld(kOpp[q], -cx, -cy, -cz);
                if (WM) {
                    const C cuw = C(cx) * uwx + C(cy) * uwy + C(cz) * uwz;
                    H[q] = fmac(K * C(6) * wOf<C>(q), cuw, H[q]);
                }
            } else if (nf == kSlipTop || nf == kSlipBottom) {
                constexpr int m = kMirY[q];
                if (FROM_L)
                    H[q] = pop.ld(kOpp[m], -cx, 0, -cz);
                else
                    H[q] = pop.ld(m, 0, -cy, 0);
            } else if (nf == kSlipFront || nf == kSlipBack) {
                constexpr int m = kMirZ[q];
                if (FROM_L)
                    H[q] = pop.ld(kOpp[m], -cx, -cy, 0);
                else
                    H[q] = pop.ld(m, 0, 0, -cz);
            }
        });
    }

    const Macro<C> mac = collide<C>(H, K, C(a.tau0), C(a.lam), C(a.pre));

    staticFor<kQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const typename S::Bits b = storeOut(qc, H[q], false);
        if constexpr (KEEP && kNarrow) H[q] = S::fromBits(b);
    });
    if (k & kKindFeed) {  // the outlet's copy, rounded exactly as the slot store (re-rounding a
                          // stored value, as KEEP leaves it, returns the same bits)
        T* dstFeed = static_cast<T*>(a.feed) + newLayout * kQ * plane;
        if constexpr (kNarrow) {
#pragma unroll
            for (int q = 0; q < kQ; ++q)
                reinterpret_cast<typename S::Bits*>(dstFeed)[q * plane + yz] =
                    roundedBits<S>(H[q], sto, roundBits(rh, q));
        } else {
#pragma unroll
            for (int q = 0; q < kQ; ++q) S::store(dstFeed + q * plane + yz, H[q]);
        }
    }
    if (a.writeMacro) {
        a.rho[c] = static_cast<float>(mac.rho);
        a.u[c] = static_cast<float>(mac.ux);
        a.v[c] = static_cast<float>(mac.uy);
        a.w[c] = static_cast<float>(mac.uz);
    }
}

// The loads of a cell after its kind byte: a fluid cell pulls its 19
// populations; an outlet cell loads the 19 values the x-1 column published
// in the previous step (the zero-gradient copy's source). Both go into H and
// are issued before any wait, so a wave holding an outlet lane waits for
// one round trip, not two.
template <class S, template <class> class M, bool FROM_L>
__device__ inline void loadCell(const KArgs& a, const RowPos& p, const M<S>& pop, std::uint8_t k,
                                typename S::C (&H)[kQ]);


template <class S, template <class> class M, bool FROM_L>
__device__ inline void loadCell(const KArgs& a, const RowPos& p, const M<S>& pop, std::uint8_t k,
                                typename S::C (&H)[kQ]) {
    const std::uint8_t flag = k & kKindFlagMask;
    if (flag == kFluid) {
        staticFor<kQ>([&](auto qc) {
            using L = PullSlot<FROM_L, decltype(qc)::value>;
            H[decltype(qc)::value] = pop.ld(L::dir, L::ex, L::ey, L::ez);
        });
    } else if (flag == kOutlet) {
        using T = typename S::T;
        const long long plane = static_cast<long long>(a.ny) * a.nz;
        const T* src = static_cast<const T*>(a.feed) + (FROM_L ? 0 : 1) * kQ * plane + (p.y + a.ny * p.z);
#pragma unroll
        for (int q = 0; q < kQ; ++q) H[q] = S::load(src + q * plane);
    }
}

template <class S, template <class> class M, bool FROM_L, bool WM, bool PX, bool PY>
__device__ inline void stepCell(const KArgs& a, const RowPos& p) {
    using C = typename S::C;
    const M<S> pop(a, p);
    const std::uint8_t k = a.kind[nb(p, 0, 0, 0)];
    C H[kQ];
    loadCell<S, M, FROM_L>(a, p, pop, k, H);
    cellBody<S, M, FROM_L, WM, PX, PY>(a, p, pop, k, H);
}

// Resident waves per SIMD a step kernel asks the register allocator for.
// Seven (at most 72 VGPRs) for the 16-bit storages' wall-model steps and
// record-writing steps: they then fit (the L->S wall-model step spills two
// dwords; 73-107 VGPRs and 4-6 waves without the request) and run faster on
// the refined wing. The
// other variants keep the allocator's own choice (a request for one wave is
// the default): the same cap spills the 32-bit and plain-wall variants and
// slows them (both measured on MI355X).
template <class S, bool WM_OR_RECORD>
inline constexpr int kStepWaves = sizeof(typename S::T) == 2 && WM_OR_RECORD ? 7 : 1;

template <class S, template <class> class M, bool FROM_L, bool WM, bool PX, bool PY>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kStepWaves<S, WM>, 8)))
stepKernel(KArgs a) {
    RowPos p;
    if (!rowPos<PX, PY>(a, p)) return;
    stepCell<S, M, FROM_L, WM, PX, PY>(a, p);
}

// The buffer-addressed step with the register budget capped at 64 VGPRs so
// eight waves fit per SIMD (the buffer accessor keeps every variant at 43-55
// VGPRs with no spills; the extra resident waves give +4 % at FP16 on the
// 23.6 M-cell wing, measured on MI355X).
template <class S, bool FROM_L, bool WM, bool PX, bool PY>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8, 8)))
stepKernelBuf(KArgs a) {
    RowPos p;
    if (!rowPos<PX, PY>(a, p)) return;
    stepCell<S, PopBuf, FROM_L, WM, PX, PY>(a, p);
}

// ---------------------------------------------------------------------------
// Fine step that also writes the restriction's child-average records (factor
// 2, the last fine sub-step of a coarse step). The four waves of a workgroup
// take the same 64-cell segment of the rows (y, y+1) x (z, z+1), so its 256
// cells are the children of 32 coarse cells. After the ordinary step every
// wave leaves its cells' stored post-collision deviations in LDS (through the
// storage codec, as a read of the fine array returns them), and wave 0 forms
// the records exactly as the child-average pass does from the fine array:
// children summed in (z, y) order per fine column, the column pair combined
// by a cross-lane read, then the shared record arithmetic. The records are
// therefore bit-identical to that pass, which this launch replaces.
// ---------------------------------------------------------------------------
struct RecArgs {
    float4* rec;  // nR records (full or compact format, by storage)
    long long nR;
    int cx0, cx1;  // restricted coarse columns, relative to the box
    int cy0, cy1;  // restricted coarse rows, relative to the box
    int rw, rh;
};


template <class S, template <class> class M, bool FROM_L, bool WM>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kStepWaves<S, true>, 8)))
stepRecordKernel(KArgs a, RecArgs rg) {
    static_assert(kWavesPerBlock == 4, "one wave per child row of a 2 x 2 row group");
    __shared__ float sH[kWavesPerBlock][kQ][kWave];
    __shared__ unsigned char sLive[kWavesPerBlock][kWave];
    using C = typename S::C;
    const int hb = blockIdx.x;
    const int lb = (hb & (kXcds - 1)) * a.chunk + (hb >> 3);
    if (lb >= a.nblocks) return;  // workgroup-uniform
    const int wv = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) >> 6);
    const int lane = static_cast<int>(threadIdx.x) & (kWave - 1);
    const int seg = lb % a.nseg;
    const int pr = lb / a.nseg;
    const int yp = pr % (a.ny >> 1), zp = pr / (a.ny >> 1);
    const int gw = ((2 * zp + (wv >> 1)) * a.ny + 2 * yp + (wv & 1)) * a.nseg + seg;
    RowPos p;
    const bool live = taskPos<false, false>(a, gw, p);
    C H[kQ];
    std::uint8_t k = kSolid;
    if (live) {
        const M<S> pop(a, p);
        k = a.kind[nb(p, 0, 0, 0)];
        loadCell<S, M, FROM_L>(a, p, pop, k, H);
        cellBody<S, M, FROM_L, WM, false, false, true>(a, p, pop, k, H);
    }
    if (yp < rg.cy0 || yp >= rg.cy1) return;  // workgroup-uniform: rows outside the restricted box

    // H holds what the step stored (the 16-bit formats' rounded values; FP32
    // and FP64 store H exactly), which the child-average pass would read back
    const bool fluid = live && (k & kKindFlagMask) == kFluid;
#pragma unroll
    for (int q = 0; q < kQ; ++q) sH[wv][q][lane] = fluid ? static_cast<float>(H[q]) / S::K : 0.0f;
    sLive[wv][lane] = (live && (k & kKindFlagMask) != kSolid) ? 1 : 0;
    __syncthreads();
    if (wv != 0) return;

    const int lx = lane & 1;
    float h[kQ];
#pragma unroll
    for (int q = 0; q < kQ; ++q) h[q] = 0.0f;
    int nf = 0;
#pragma unroll
    for (int w = 0; w < kWavesPerBlock; ++w) {  // child (dz, dy) = (w >> 1, w & 1)
        if (!sLive[w][lane]) continue;
        ++nf;
#pragma unroll
        for (int q = 0; q < kQ; ++q) h[q] += sH[w][q][lane];
    }
    const int src = min(lane + 1, kWave - 1);
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        const float o = __shfl(h[q], src, kWave);
        if (lx == 0) h[q] += o;
    }
    const int on = __shfl(nf, src, kWave);
    if (lx == 0) nf += on;
    const int cx = (seg * kWave + lane) >> 1;
    if (lx != 0 || cx < rg.cx0 || cx >= rg.cx1) return;
    const long long ri =
        (cx - rg.cx0) + static_cast<long long>(rg.rw) * ((yp - rg.cy0) + static_cast<long long>(rg.rh) * zp);
    float rec[4 * kRecordPlanes];
    childRecord(h, nf, rec);
    storeRecord<RecordFormat<S>::kCompact>(rg.rec, rg.nR, ri, rec);
}

// ---------------------------------------------------------------------------
// Batched step of a group of independent lattices (e.g. the cases of a polar
// sweep, each with its own grid size, layout parity, relaxation time and
// inlet speed) in one launch. The row-segment tasks of all members are
// concatenated; a wave finds its member from the task bounds (kernel
// arguments), reads that member's argument record with scalar loads and runs
// the same cell update as the single-lattice kernel. Launch geometry (waves
// per workgroup, XCD slab or linear task order) is a tuning choice that does
// not change the result.
// ---------------------------------------------------------------------------
struct GroupEntry {
    KArgs a;
    int fromL;  // member array in layout L this step
    int wm;     // member has the wall-model slip field
    int taskBase;
    int pad;
};

struct GroupLaunch {
    int members;
    int totalTasks;
    int wavesPerBlock;
    int slabOrder;  // 1: each XCD takes a contiguous slab of the task range
    int nblocks;
    int chunk;      // workgroups per XCD slab
    int bound[kGroupMaxMembers + 1];  // task prefix: member m owns [bound[m], bound[m+1])
};

// Member and local task of this wave; false past the end of the launch.
__device__ inline bool groupTask(const GroupLaunch& g, int& m, int& local) {
    const int hb = blockIdx.x;
    const int lb = g.slabOrder ? (hb & (kXcds - 1)) * g.chunk + (hb >> 3) : hb;
    if (lb >= g.nblocks) return false;
    const int wv = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) >> 6);
    const int gw = lb * g.wavesPerBlock + wv;
    if (gw >= g.totalTasks) return false;
    m = 0;
    while (m + 1 < g.members && gw >= g.bound[m + 1]) ++m;
    local = gw - g.bound[m];
    return true;
}

// Launch whose members share the layout parity and the wall-model setting
// (the usual case: members stepped together from their initial state).
template <class S, bool FROM_L, bool WM>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(8, 8)))
stepKernelGroup(const GroupEntry* __restrict__ tab, GroupLaunch g) {
    int m, local;
    if (!groupTask(g, m, local)) return;
    const KArgs& a = tab[m].a;
    RowPos p;
    if (!taskPos<false, false>(a, local, p)) return;
    stepCell<S, PopBuf, FROM_L, WM, false, false>(a, p);
}

// Members with mixed parity or wall-model settings: per-wave dispatch.
template <class S>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(8, 8)))
stepKernelGroupMixed(const GroupEntry* __restrict__ tab, GroupLaunch g) {
    int m, local;
    if (!groupTask(g, m, local)) return;
    const GroupEntry& e = tab[m];
    const KArgs& a = e.a;
    RowPos p;
    if (!taskPos<false, false>(a, local, p)) return;
    if (e.fromL) {
        if (e.wm) stepCell<S, PopBuf, true, true, false, false>(a, p);
        else stepCell<S, PopBuf, true, false, false, false>(a, p);
    } else {
        if (e.wm) stepCell<S, PopBuf, false, true, false, false>(a, p);
        else stepCell<S, PopBuf, false, false, false, false>(a, p);
    }
}

// ---------------------------------------------------------------------------
// Init: f = feq(1, (u0,0,0)) (solid cells at rest) written in layout L, plus
// both halves of the outlet feed planes (reference initEquilibriumKernel,
// lbm_core.cu:102-118, run on both ping-pong buffers, lbm_solver.cpp:931-935):
// a feed cell that never collides keeps this state in both halves.
// ---------------------------------------------------------------------------
template <class S>
__global__ void __launch_bounds__(kBlock) initKernel(KArgs a, float u0) {
    using T = typename S::T;
    using C = typename S::C;
    RowPos p;
    if (!rowPos<false, false>(a, p)) return;
    T* const f = static_cast<T*>(a.f);
    const unsigned c = nb(p, 0, 0, 0);
    const std::uint8_t k = a.kind[c];
    const C ux = ((k & kKindFlagMask) == kSolid) ? C(0) : C(u0);
    const C K = C(S::K);
    const C usq = C(1.5) * ux * ux;
    const long long plane = static_cast<long long>(a.ny) * a.nz;
    const unsigned yz = static_cast<unsigned>(p.y + a.ny * p.z);
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        const C h = heqOf<C>(q, K, C(0), C(1), ux, C(0), C(0), usq);
        S::store(f + static_cast<long long>(kOpp[q]) * a.ncells + c, h);
        if (k & kKindFeed) {
            T* fd = static_cast<T*>(a.feed);
            S::store(fd + q * plane + yz, h);
            S::store(fd + (kQ + q) * plane + yz, h);
        }
    }
}

// Spanwise speck perturbation on fluid cells, layout L (reference
// spanwisePerturbKernel, lbm_core.cu:130-160, same fast sine intrinsic).
template <class S>
__global__ void __launch_bounds__(kBlock) speckKernel(KArgs a, float amp, float ph1, float ph2,
                                                      float ph3) {
    using T = typename S::T;
    using C = typename S::C;
    RowPos p;
    if (!rowPos<false, false>(a, p)) return;
    T* const f = static_cast<T*>(a.f);
    const unsigned c = nb(p, 0, 0, 0);
    if ((a.kind[c] & kKindFlagMask) != kFluid) return;
    const float tz = 6.2831853f * (static_cast<float>(p.z) + 0.5f) / a.nz;
    const float tx = 6.2831853f * (static_cast<float>(p.x) + 0.5f) / a.nx;
    const float wz = amp * (0.6f * __sinf(3.0f * tz + ph1) + 0.4f * __sinf(7.0f * tz + ph2)) *
                     (1.0f + 0.3f * __sinf(2.0f * tx + ph3));
    const C K = C(S::K);
    // A cell that feeds the outlet also refreshes both feed-plane copies, so
    // the first outlet copy sees the perturbed column as the reference's
    // direct x-1 read does.
    const bool feed = (a.kind[c] & kKindFeed) != 0;
    T* const fd = static_cast<T*>(a.feed);
    const long long plane = static_cast<long long>(a.ny) * a.nz;
    const unsigned yz = static_cast<unsigned>(p.y + a.ny * p.z);
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        if (kCz[q] == 0) continue;
        T* slot = f + static_cast<long long>(kOpp[q]) * a.ncells + c;
        const C h = S::load(slot) + K * C(3) * wOf<C>(q) * C(kCz[q]) * C(wz);
        S::store(slot, h);
        if (feed) {
            S::store(fd + q * plane + yz, h);
            S::store(fd + (kQ + q) * plane + yz, h);
        }
    }
}

// Canonical unpadded SoA (FP32 or FP64) -> layout L, and the half of the
// outlet feed planes the next step reads. The other half keeps what it held,
// as the reference's uploadF replaces only the active ping-pong buffer
// (tests/lbm_test_util.h:116-147): a feed cell that never collides (a slip
// cell beside an outlet corner) then still alternates between the uploaded
// state and the one before it, like the reference's two buffers.
template <class S, class CT>
__global__ void __launch_bounds__(kBlock) uploadKernel(KArgs a, const CT* canon) {
    using T = typename S::T;
    using C = typename S::C;
    RowPos p;
    if (!rowPos<false, false>(a, p)) return;
    T* const f = static_cast<T*>(a.f);
    const unsigned c = nb(p, 0, 0, 0);
    const std::uint8_t k = a.kind[c];
    const C K = C(S::K);
    const long long plane = static_cast<long long>(a.ny) * a.nz;
    const unsigned yz = static_cast<unsigned>(p.y + a.ny * p.z);
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        const C h = K * (C(canon[static_cast<long long>(q) * a.ncells + c]) - wOf<C>(q));
        S::store(f + static_cast<long long>(kOpp[q]) * a.ncells + c, h);
        if (k & kKindFeed) S::store(static_cast<T*>(a.feed) + q * plane + yz, h);  // read from layout L
    }
}

// Layout L or S -> canonical SoA. Slots of layout S that fall outside a
// non-periodic domain are unrecoverable and come back as NaN.
template <class S, class CT, bool PX, bool PY>
__global__ void __launch_bounds__(kBlock) downloadKernel(KArgs a, int layout, CT* canon) {
    using T = typename S::T;
    using C = typename S::C;
    RowPos p;
    if (!rowPos<PX, PY>(a, p)) return;
    const T* const f = static_cast<const T*>(a.f);
    const unsigned c = nb(p, 0, 0, 0);
    const C invK = C(1) / C(S::K);
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        CT out;
        if (layout == kLayoutL) {
            out = CT(wOf<C>(q) + S::load(f + static_cast<long long>(kOpp[q]) * a.ncells + c) * invK);
        } else if (inDomain<PX>(a, p, kCx[q], kCy[q])) {
            out = CT(wOf<C>(q) +
                     S::load(f + static_cast<long long>(q) * a.ncells + nb(p, kCx[q], kCy[q], kCz[q])) *
                         invK);
        } else {
            out = CT(__builtin_nan(""));
        }
        canon[static_cast<long long>(q) * a.ncells + c] = out;
    }
}

// Compact-snapshot restore (reference initFromMacroKernel, lbm_core.cu:162-198):
// f = feq(rho, u) from the macro arrays, solid cells at rest, written in
// layout L plus both halves of the outlet feed planes.
template <class S>
__global__ void __launch_bounds__(kBlock) macroInitKernel(KArgs a) {
    using T = typename S::T;
    using C = typename S::C;
    RowPos p;
    if (!rowPos<false, false>(a, p)) return;
    T* const f = static_cast<T*>(a.f);
    const unsigned c = nb(p, 0, 0, 0);
    const std::uint8_t k = a.kind[c];
    const bool solid = (k & kKindFlagMask) == kSolid;
    const C r = solid ? C(1) : C(a.rho[c]);
    const C ux = solid ? C(0) : C(a.u[c]);
    const C uy = solid ? C(0) : C(a.v[c]);
    const C uz = solid ? C(0) : C(a.w[c]);
    const C K = C(S::K);
    const C usq = C(1.5) * (ux * ux + uy * uy + uz * uz);
    const long long plane = static_cast<long long>(a.ny) * a.nz;
    const unsigned yz = static_cast<unsigned>(p.y + a.ny * p.z);
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        const C h = heqOf<C>(q, K, r - C(1), r, ux, uy, uz, usq);
        S::store(f + static_cast<long long>(kOpp[q]) * a.ncells + c, h);
        if (k & kKindFeed) {
            T* fd = static_cast<T*>(a.feed);
            S::store(fd + q * plane + yz, h);
            S::store(fd + (kQ + q) * plane + yz, h);
        }
    }
}

// Cell index of (x + ex, y + ey, z + ez): z wraps, x / y wrap when periodic,
// -1 when the position leaves a non-periodic x / y range.
__device__ inline long long shiftedCell(const KArgs& a, int x, int y, int z, int ex, int ey, int ez,
                                        bool px, bool py) {
    int xx = x + ex, yy = y + ey, zz = z + ez;
    if (xx < 0 || xx >= a.nx) {
        if (!px) return -1;
        xx = xx < 0 ? xx + a.nx : xx - a.nx;
    }
    if (yy < 0 || yy >= a.ny) {
        if (!py) return -1;
        yy = yy < 0 ? yy + a.ny : yy - a.ny;
    }
    zz = zz < 0 ? zz + a.nz : (zz >= a.nz ? zz - a.nz : zz);
    return xx + static_cast<long long>(a.nx) * (yy + static_cast<long long>(a.ny) * zz);
}

// Storage slot of population q of cell (x, y, z) in layout L (in place,
// direction reversed) or S (at the downstream neighbour), -1 when layout S
// puts it outside a non-periodic domain.
__device__ inline long long popSlot(const KArgs& a, int layout, int q, int x, int y, int z, bool px,
                                    bool py) {
    if (layout == kLayoutL)
        return static_cast<long long>(kOpp[q]) * a.ncells + x +
               static_cast<long long>(a.nx) * (y + static_cast<long long>(a.ny) * z);
    const long long n = shiftedCell(a, x, y, z, kCx[q], kCy[q], kCz[q], px, py);
    return n < 0 ? -1 : static_cast<long long>(q) * a.ncells + n;
}

// Warm-restart flag-edit fixup (reference applyFlagEditsKernel,
// lbm_core.cu:644-717), one thread per edited cell of a host-built list. A
// newly solid cell gets f = 0; a newly fluid cell the equilibrium at the mean
// of rho and u = j / rho over its valid neighbours (rest state when none).
// Every neighbour read is an unedited cell and every write a slot of an
// edited one, so the threads never touch the same slot.
template <class S>
__global__ void __launch_bounds__(kBlock) flagEditKernel(KArgs a, int layout, int px, int py,
                                                         const FlagEdit* edits, int count) {
    using T = typename S::T;
    using C = typename S::C;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= count) return;
    const FlagEdit e = edits[i];
    if (e.cell < 0 || e.cell >= a.ncells) return;
    T* const f = static_cast<T*>(a.f);
    const long long nxny = static_cast<long long>(a.nx) * a.ny;
    const int z = static_cast<int>(e.cell / nxny);
    const long long rem = e.cell - z * nxny;
    const int y = static_cast<int>(rem / a.nx);
    const int x = static_cast<int>(rem - static_cast<long long>(y) * a.nx);
    const C K = C(S::K);
    const C invK = C(1) / K;
    C hOut[kQ];
    if (e.kind == 1) {
#pragma unroll
        for (int q = 0; q < kQ; ++q) hOut[q] = -K * wOf<C>(q);
    } else {
        C rho = 0, ux = 0, uy = 0, uz = 0;
        int found = 0;
        for (int q = 1; q < kQ; ++q) {
            if (!(e.nbMask & (1u << q))) continue;
            const long long n = shiftedCell(a, x, y, z, kCx[q], kCy[q], kCz[q], false, false);
            if (n < 0) continue;
            const int nz = static_cast<int>(n / nxny);
            const long long nr = n - nz * nxny;
            const int ny = static_cast<int>(nr / a.nx);
            const int nx = static_cast<int>(nr - static_cast<long long>(ny) * a.nx);
            C r = 0, jx = 0, jy = 0, jz = 0;
            bool ok = true;
#pragma unroll
            for (int p = 0; p < kQ; ++p) {
                const long long s = popSlot(a, layout, p, nx, ny, nz, px != 0, py != 0);
                if (s < 0) {
                    ok = false;
                    continue;
                }
                const C fq = wOf<C>(p) + S::load(f + s) * invK;
                r += fq;
                jx += C(kCx[p]) * fq;
                jy += C(kCy[p]) * fq;
                jz += C(kCz[p]) * fq;
            }
            if (!ok) continue;
            rho += r;
            ux += jx / r;
            uy += jy / r;
            uz += jz / r;
            ++found;
        }
        if (found > 0) {
            const C inv = C(1) / C(found);
            rho *= inv;
            ux *= inv;
            uy *= inv;
            uz *= inv;
        } else {
            rho = C(1);
            ux = uy = uz = C(0);
        }
        const C usq = C(1.5) * (ux * ux + uy * uy + uz * uz);
#pragma unroll
        for (int q = 0; q < kQ; ++q) hOut[q] = heqOf<C>(q, K, rho - C(1), rho, ux, uy, uz, usq);
    }
    const bool feed = (a.kind[e.cell] & kKindFeed) != 0;
    T* const fd = static_cast<T*>(a.feed);
    const long long plane = static_cast<long long>(a.ny) * a.nz;
    const long long yz = y + static_cast<long long>(a.ny) * z;
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        const long long s = popSlot(a, layout, q, x, y, z, px != 0, py != 0);
        if (s >= 0) S::store(f + s, hOut[q]);
        if (feed) {
            S::store(fd + q * plane + yz, hOut[q]);
            S::store(fd + (kQ + q) * plane + yz, hOut[q]);
        }
    }
}

// Momentum exchange (reference forceReductionKernel<WM>, lbm_core.cu:573-637):
// F = sum over fluid->solid links of c_q (2 f*_q - 6 w_q c_q.u_w), per-link
// terms in the compute type, sums in FP64 in a fixed order.
template <class S>
__global__ void __launch_bounds__(kBlock) forceKernel(KArgs a, int layout, const int* cells,
                                                      const std::uint32_t* masks, int count,
                                                      double* partials) {
    using C = typename S::C;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double acc[3] = {0, 0, 0};
    if (i < count) {
        const int cell = cells[i];
        const std::uint32_t mask = masks[i];
        const int z = cell / (a.nx * a.ny);
        const int rem = cell - z * a.nx * a.ny;
        const int y = rem / a.nx;
        const int x = rem - y * a.nx;
        C uwx = 0, uwy = 0, uwz = 0;
        if (a.uwx) {
            uwx = C(halfBits(a.uwx[cell]));
            uwy = C(halfBits(a.uwy[cell]));
            uwz = C(halfBits(a.uwz[cell]));
        }
        const C invK = C(1) / C(S::K);
        C fx = 0, fy = 0, fz = 0;
#pragma unroll
        for (int q = 1; q < kQ; ++q) {
            if (!(mask & (1u << q))) continue;
            const C fq = wOf<C>(q) + postH<S>(a, layout, q, x, y, z) * invK;
            const C corr = C(6) * wOf<C>(q) * (C(kCx[q]) * uwx + C(kCy[q]) * uwy + C(kCz[q]) * uwz);
            const C m = C(2) * fq - corr;
            fx += C(kCx[q]) * m;
            fy += C(kCy[q]) * m;
            fz += C(kCz[q]) * m;
        }
        acc[0] = static_cast<double>(fx);
        acc[1] = static_cast<double>(fy);
        acc[2] = static_cast<double>(fz);
    }
    blockSum<3>(acc, partials + 3 * blockIdx.x);
}

// Fixed-order final sum of NV-wide partials by one block.
template <int NV>
__global__ void __launch_bounds__(kBlock) finalSumKernel(const double* partials, int slots,
                                                         double* out) {
    double acc[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) acc[i] = 0;
    for (int s = threadIdx.x; s < slots; s += kBlock)
#pragma unroll
        for (int i = 0; i < NV; ++i) acc[i] += partials[NV * s + i];
    blockSum<NV>(acc, out);
}

// Global mass sum_{fluid cells} sum_q f*_q in FP64 (grid-stride, fixed order).
template <class S>
__global__ void __launch_bounds__(kBlock) massKernel(KArgs a, int layout, double* partials) {
    using C = typename S::C;
    const C invK = C(1) / C(S::K);
    double acc[1] = {0};
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    for (long long cell = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; cell < a.ncells;
         cell += stride) {
        if ((a.kind[cell] & kKindFlagMask) != kFluid) continue;
        const int z = static_cast<int>(cell / (static_cast<long long>(a.nx) * a.ny));
        const int rem = static_cast<int>(cell - static_cast<long long>(z) * a.nx * a.ny);
        const int y = rem / a.nx;
        const int x = rem - y * a.nx;
        C s = 0;
#pragma unroll
        for (int q = 0; q < kQ; ++q) s += postH<S>(a, layout, q, x, y, z);
        acc[0] += 1.0 + static_cast<double>(s * invK);
    }
    blockSum<1>(acc, partials + blockIdx.x);
}

// NaN watchdog (reference nanWatchdogKernel, lbm_core.cu:522-559).
template <class S>
__global__ void __launch_bounds__(kBlock) watchdogKernel(KArgs a, int layout, long long offset,
                                                         int flagBase, int* flag) {
    const long long t = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x;
    const long long cell = t * 4093 + offset;
    if (cell >= a.ncells) return;
    // a refined level's shell holds mixed-time slots that are never read
    if ((a.kind[cell] & kKindFlagMask) == kInterface) return;
    const int z = static_cast<int>(cell / (static_cast<long long>(a.nx) * a.ny));
    const int rem = static_cast<int>(cell - static_cast<long long>(z) * a.nx * a.ny);
    const int y = rem / a.nx;
    const int x = rem - y * a.nx;
    const float invK = 1.0f / static_cast<float>(S::K);
    float s = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        const float fq = wOf<float>(q) + static_cast<float>(postH<S>(a, layout, q, x, y, z)) * invK;
        s += fq;
        jx += kCx[q] * fq;
        jy += kCy[q] * fq;
        jz += kCz[q] * fq;
    }
    if (!__builtin_isfinite(s)) {
        *flag = flagBase + 1;
    } else {
        const float speed2 = (jx * jx + jy * jy + jz * jz) / fmaxf(s * s, 1e-12f);
        const float cap = 0.95f * 0.30f;
        if (speed2 > cap * cap) *flag = flagBase + 2;
    }
}

__global__ void fillFloatKernel(float* dst, long long n, float value) {
    const long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
    if (i < n) dst[i] = value;
}

unsigned rowGrid(const KArgs& a) { return static_cast<unsigned>(a.chunk * kXcds); }

int lastError() { return static_cast<int>(hipGetLastError()); }

// Launch of one step kernel family K<FROM_L, WM, PX, PY> by layout, wall
// model and periodicity (periodic x / y are only used together: the fully
// periodic box).
template <template <bool, bool, bool, bool> class K>
int dispatchStepK(const KArgs& a, const LatticeView& lat, int layout, hipStream_t st) {
    const bool wm = lat.uwx && lat.uwy && lat.uwz;
    const bool fromL = layout == kLayoutL;
    const bool pxy = lat.periodicX || lat.periodicY;
    const dim3 g(rowGrid(a)), b(kBlock);
    if (pxy) {
        if (fromL) K<true, false, true, true>::launch(g, b, st, a);
        else K<false, false, true, true>::launch(g, b, st, a);
    } else if (wm) {
        if (fromL) K<true, true, false, false>::launch(g, b, st, a);
        else K<false, true, false, false>::launch(g, b, st, a);
    } else {
        if (fromL) K<true, false, false, false>::launch(g, b, st, a);
        else K<false, false, false, false>::launch(g, b, st, a);
    }
    return lastError();
}

template <class S>
struct PtrStep {
    template <bool FROM_L, bool WM, bool PX, bool PY>
    struct K {
        static void launch(dim3 g, dim3 b, hipStream_t st, const KArgs& a) {
            stepKernel<S, PopPtr, FROM_L, WM, PX, PY><<<g, b, 0, st>>>(a);
        }
    };
};

template <class S>
struct BufStep {
    template <bool FROM_L, bool WM, bool PX, bool PY>
    struct K {
        static void launch(dim3 g, dim3 b, hipStream_t st, const KArgs& a) {
            stepKernelBuf<S, FROM_L, WM, PX, PY><<<g, b, 0, st>>>(a);
        }
    };
};

// The default one-cell kernel uses the flat-pointer accessor: on the 23.6 M-
// cell wing it ran 10 % faster than the buffer-addressed kernel at 16 bits
// (0.485 vs 0.539 ms) and level at FP32. The buffer-addressed kernel (kStepBuf)
// needs every population byte offset in 31 bits (FP16 / BF16 up to 56 M
// cells, FP32 up to 28 M) and falls back to the default beyond; FP64 always
// takes the default.
template <class S>
int dispatchStep(const KArgs& a, const LatticeView& lat, int layout, int variant, hipStream_t st) {
    const unsigned long long bytes =
        static_cast<unsigned long long>(kQ) * static_cast<unsigned long long>(a.ncells) * sizeof(typename S::T);
    if constexpr (!std::is_same_v<S, StoreF64>) {
        if (variant == kStepBuf && bytes < (1ull << 31))
            return dispatchStepK<BufStep<S>::template K>(a, lat, layout, st);
    }
    return dispatchStepK<PtrStep<S>::template K>(a, lat, layout, st);
}

template <template <class> class F, class... Args>
int byStorage(int storage, Args&&... args) {
    switch (storage) {
        case kStoreF32: return F<StoreF32>::run(args...);
        case kStoreF16: return F<StoreF16>::run(args...);
        case kStoreBF16: return F<StoreBF16>::run(args...);
        case kStoreF16C: return F<StoreF16C>::run(args...);
        case kStoreF64: return F<StoreF64>::run(args...);
        default: return static_cast<int>(hipErrorInvalidValue);
    }
}

template <class S>
struct StepOp {
    static int run(const LatticeView& lat, int layout, const StepArgs& sa, hipStream_t st) {
        KArgs a = makeArgs(lat);
        a.tau0 = sa.tau0;
        a.lam = sa.magicLambda;
        a.pre = sa.smagPrefactor;
        a.uIn = sa.uInlet;
        a.writeMacro = (sa.writeMacro && lat.rho && lat.u && lat.v && lat.w) ? 1 : 0;
        if (sa.variant == kStepWide) a.wideIndex = 1;
        a.stochastic = sa.stochastic;
        a.roundSeed = sa.roundSeed;
        return dispatchStep<S>(a, lat, layout, sa.variant, st);
    }
};

template <class S>
struct StepRecordOp {
    static int run(const LatticeView& lat, int layout, const StepArgs& sa, const RecArgs& rg, hipStream_t st) {
        KArgs a = makeArgs(lat);
        a.tau0 = sa.tau0;
        a.lam = sa.magicLambda;
        a.pre = sa.smagPrefactor;
        a.uIn = sa.uInlet;
        a.writeMacro = (sa.writeMacro && lat.rho && lat.u && lat.v && lat.w) ? 1 : 0;
        if (sa.variant == kStepWide) a.wideIndex = 1;
        a.stochastic = sa.stochastic;
        a.roundSeed = sa.roundSeed;
        const bool wm = lat.uwx && lat.uwy && lat.uwz;
        const bool fromL = layout == kLayoutL;
        const dim3 g(rowGrid(a)), b(kBlock);
        if (fromL && wm) launch<true, true>(g, b, st, a, rg);
        else if (fromL) launch<true, false>(g, b, st, a, rg);
        else if (wm) launch<false, true>(g, b, st, a, rg);
        else launch<false, false>(g, b, st, a, rg);
        return lastError();
    }
    template <bool FROM_L, bool WM>
    static void launch(dim3 g, dim3 b, hipStream_t st, const KArgs& a, const RecArgs& rg) {
        stepRecordKernel<S, PopPtr, FROM_L, WM><<<g, b, 0, st>>>(a, rg);
    }
};

template <class S>
struct InitOp {
    static int run(const LatticeView& lat, float u0, hipStream_t st) {
        KArgs a = makeArgs(lat);
        initKernel<S><<<rowGrid(a), kBlock, 0, st>>>(a, u0);
        return lastError();
    }
};

template <class S>
struct SpeckOp {
    static int run(const LatticeView& lat, float amp, float p1, float p2, float p3, hipStream_t st) {
        KArgs a = makeArgs(lat);
        speckKernel<S><<<rowGrid(a), kBlock, 0, st>>>(a, amp, p1, p2, p3);
        return lastError();
    }
};

template <class S>
struct UploadOp {
    static int run(const LatticeView& lat, const void* canon, int isF64, hipStream_t st) {
        KArgs a = makeArgs(lat);
        if (isF64)
            uploadKernel<S, double><<<rowGrid(a), kBlock, 0, st>>>(a, static_cast<const double*>(canon));
        else
            uploadKernel<S, float><<<rowGrid(a), kBlock, 0, st>>>(a, static_cast<const float*>(canon));
        return lastError();
    }
};

template <class S>
struct DownloadOp {
    static int run(const LatticeView& lat, int layout, void* canon, int isF64, hipStream_t st) {
        KArgs a = makeArgs(lat);
        const bool pxy = lat.periodicX || lat.periodicY;
        if (isF64) {
            if (pxy)
                downloadKernel<S, double, true, true><<<rowGrid(a), kBlock, 0, st>>>(a, layout, static_cast<double*>(canon));
            else
                downloadKernel<S, double, false, false><<<rowGrid(a), kBlock, 0, st>>>(a, layout, static_cast<double*>(canon));
        } else {
            if (pxy)
                downloadKernel<S, float, true, true><<<rowGrid(a), kBlock, 0, st>>>(a, layout, static_cast<float*>(canon));
            else
                downloadKernel<S, float, false, false><<<rowGrid(a), kBlock, 0, st>>>(a, layout, static_cast<float*>(canon));
        }
        return lastError();
    }
};

template <class S>
struct MacroInitOp {
    static int run(const LatticeView& lat, hipStream_t st) {
        KArgs a = makeArgs(lat);
        if (!a.rho || !a.u || !a.v || !a.w) return static_cast<int>(hipErrorInvalidValue);
        macroInitKernel<S><<<rowGrid(a), kBlock, 0, st>>>(a);
        return lastError();
    }
};

template <class S>
struct FlagEditOp {
    static int run(const LatticeView& lat, int layout, const FlagEdit* edits, int count, hipStream_t st) {
        KArgs a = makeArgs(lat);
        if (count <= 0) return 0;
        const int blocks = (count + kBlock - 1) / kBlock;
        flagEditKernel<S><<<blocks, kBlock, 0, st>>>(a, layout, lat.periodicX, lat.periodicY, edits, count);
        return lastError();
    }
};

template <class S>
struct ForceOp {
    static int run(const LatticeView& lat, int layout, const int* cells, const std::uint32_t* masks,
                   int count, double* partials, int slots, double* out3, hipStream_t st) {
        KArgs a = makeArgs(lat);
        const int blocks = (count + kBlock - 1) / kBlock;
        if (blocks > slots) return static_cast<int>(hipErrorInvalidValue);
        if (blocks > 0) forceKernel<S><<<blocks, kBlock, 0, st>>>(a, layout, cells, masks, count, partials);
        finalSumKernel<3><<<1, kBlock, 0, st>>>(partials, blocks, out3);
        return lastError();
    }
};

template <class S>
struct MassOp {
    static int run(const LatticeView& lat, int layout, double* partials, int slots, double* out1,
                   hipStream_t st) {
        KArgs a = makeArgs(lat);
        massKernel<S><<<slots, kBlock, 0, st>>>(a, layout, partials);
        finalSumKernel<1><<<1, kBlock, 0, st>>>(partials, slots, out1);
        return lastError();
    }
};

template <class S>
struct WatchdogOp {
    static int run(const LatticeView& lat, int layout, long long offset, int flagBase, int* flag,
                   hipStream_t st) {
        KArgs a = makeArgs(lat);
        const long long stride = 4093;
        const long long off = ((offset % stride) + stride) % stride;
        const long long samples = (a.ncells + stride - 1) / stride;
        const unsigned blocks = static_cast<unsigned>((samples + kBlock - 1) / kBlock);
        watchdogKernel<S><<<blocks, kBlock, 0, st>>>(a, layout, off, flagBase, flag);
        return lastError();
    }
};

}  // namespace

int launchInitEquilibrium(const LatticeView& lat, float uInit, void* stream) {
    return byStorage<InitOp>(lat.storage, lat, uInit, static_cast<hipStream_t>(stream));
}

int launchSpanwisePerturb(const LatticeView& lat, float amplitude, float phase1, float phase2,
                          float phase3, void* stream) {
    return byStorage<SpeckOp>(lat.storage, lat, amplitude, phase1, phase2, phase3,
                              static_cast<hipStream_t>(stream));
}

int launchUploadCanonical(const LatticeView& lat, const void* canon, int canonIsF64, void* stream) {
    return byStorage<UploadOp>(lat.storage, lat, canon, canonIsF64, static_cast<hipStream_t>(stream));
}

int launchDownloadCanonical(const LatticeView& lat, int layout, void* canon, int canonIsF64,
                            void* stream) {
    return byStorage<DownloadOp>(lat.storage, lat, layout, canon, canonIsF64,
                                 static_cast<hipStream_t>(stream));
}

int launchInitFromMacro(const LatticeView& lat, void* stream) {
    return byStorage<MacroInitOp>(lat.storage, lat, static_cast<hipStream_t>(stream));
}

int launchFlagEdits(const LatticeView& lat, int layout, const FlagEdit* edits, int count, void* stream) {
    return byStorage<FlagEditOp>(lat.storage, lat, layout, edits, count, static_cast<hipStream_t>(stream));
}

int launchStep(const LatticeView& lat, int layout, const StepArgs& args, void* stream) {
    return byStorage<StepOp>(lat.storage, lat, layout, args, static_cast<hipStream_t>(stream));
}

int stepRecordsSupported(const LatticeView& fine, const RefineGeom& g) {
    const int w = g.box.x1 - g.box.x0, h = g.box.y1 - g.box.y0;
    return g.factor == 2 && fine.dims.nx == g.fine.nx && fine.dims.ny == g.fine.ny && fine.dims.nz == g.fine.nz &&
           g.fine.nx == 2 * w && g.fine.ny == 2 * h && g.fine.nz == 2 * g.coarse.nz && !fine.periodicX &&
           !fine.periodicY && w > 2 * kRestrictBand && h > 2 * kRestrictBand;
}

int launchStepRecords(const LatticeView& fine, int layout, const StepArgs& args, const RefineGeom& g,
                      float* scratch, void* stream) {
    if (!scratch || !stepRecordsSupported(fine, g)) return static_cast<int>(hipErrorInvalidValue);
    const int w = g.box.x1 - g.box.x0, h = g.box.y1 - g.box.y0;
    RecArgs rg{};
    rg.rec = reinterpret_cast<float4*>(scratch);
    rg.cx0 = kRestrictBand;
    rg.cx1 = w - kRestrictBand;
    rg.cy0 = kRestrictBand;
    rg.cy1 = h - kRestrictBand;
    rg.rw = rg.cx1 - rg.cx0;
    rg.rh = rg.cy1 - rg.cy0;
    rg.nR = static_cast<long long>(rg.rw) * rg.rh * g.coarse.nz;
    return byStorage<StepRecordOp>(fine.storage, fine, layout, args, rg, static_cast<hipStream_t>(stream));
}

int forcePartialSlots(int count) { return (count + kBlock - 1) / kBlock + 1; }

int launchForce(const LatticeView& lat, int layout, const int* cells, const std::uint32_t* masks,
                int count, double* partials, int partialSlots, double* out3, void* stream) {
    return byStorage<ForceOp>(lat.storage, lat, layout, cells, masks, count, partials, partialSlots,
                              out3, static_cast<hipStream_t>(stream));
}

int massPartialSlots(const Dims& d) {
    const long long blocks = (d.cells() + kBlock - 1) / kBlock;
    return static_cast<int>(blocks < 1024 ? (blocks < 1 ? 1 : blocks) : 1024);
}

int launchMass(const LatticeView& lat, int layout, double* partials, int partialSlots, double* out1,
               void* stream) {
    return byStorage<MassOp>(lat.storage, lat, layout, partials, partialSlots, out1,
                             static_cast<hipStream_t>(stream));
}

int launchWatchdog(const LatticeView& lat, int layout, long long sampleOffset, int flagBase,
                   int* flag, void* stream) {
    return byStorage<WatchdogOp>(lat.storage, lat, layout, sampleOffset, flagBase, flag,
                                 static_cast<hipStream_t>(stream));
}

// ---- allocation-free force reduction ----------------------------------------
// File-scope device scratch of the allocation-free force entry point: the
// force kernel's per-block partial sums (3 doubles per block) and the final
// sum. Every GPU holds its own copy for the life of the process; the host
// reaches it only through the address hipGetSymbolAddress returns for the
// current device (the symbol's name is not a device address in host code).
// Calls that overlap on one GPU share it and race.
static __device__ __align__(256) double scratch[3 * kNoAllocForceSlots + 4];

namespace {
std::mutex gScratchMutex;
double* gScratchPtr[kMaxScratchDevices] = {};
}  // namespace

int noAllocScratch(double** out) {
    *out = nullptr;
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxScratchDevices) {
        (void)hipGetLastError();
        return static_cast<int>(hipErrorInvalidDevice);
    }
    std::lock_guard<std::mutex> lock(gScratchMutex);
    if (!gScratchPtr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return static_cast<int>(e != hipSuccess ? e : hipErrorInvalidSymbol);
        }
        gScratchPtr[dev] = static_cast<double*>(p);
    }
    *out = gScratchPtr[dev];
    return 0;
}

int launchForceNoAlloc(const LatticeView& lat, int layout, const int* cells, const std::uint32_t* masks, int count,
                       double* host3, void* stream) {
    if (count < 0 || !host3 || (count > 0 && (!cells || !masks))) return static_cast<int>(hipErrorInvalidValue);
    if ((count + kBlock - 1) / kBlock > kNoAllocForceSlots) return static_cast<int>(hipErrorInvalidValue);
    double* d = nullptr;
    if (const int e = noAllocScratch(&d); e != 0) return e;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    double* out3 = d + 3 * kNoAllocForceSlots;
    // every partial a launch reads is written by the same launch, and the
    // final sum overwrites the result slot: nothing needs clearing
    if (const int e = byStorage<ForceOp>(lat.storage, lat, layout, cells, masks, count, d, kNoAllocForceSlots, out3,
                                         st);
        e != 0)
        return e;
    if (hipMemcpyAsync(host3, out3, 3 * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return lastError();
    return 0;
}

int launchFillFloat(float* dst, long long count, float value, void* stream) {
    if (!dst || count <= 0) return static_cast<int>(hipErrorInvalidValue);
    const unsigned blocks = static_cast<unsigned>((count + kBlock - 1) / kBlock);
    fillFloatKernel<<<blocks, kBlock, 0, static_cast<hipStream_t>(stream)>>>(dst, count, value);
    return lastError();
}


// ---- lattice groups ---------------------------------------------------------
std::size_t groupEntryBytes() { return sizeof(GroupEntry); }

int groupTasks(const LatticeView& lat) { return makeArgs(lat).nwaves; }

int groupEligible(const LatticeView& lat) {
    if (lat.storage == kStoreF64 || lat.periodicX || lat.periodicY) return 0;
    const unsigned long long bytes = static_cast<unsigned long long>(kQ) *
                                     static_cast<unsigned long long>(lat.dims.cells()) * storageBytes(lat.storage);
    return bytes < (1ull << 31) ? 1 : 0;
}

int buildGroupEntry(const LatticeView& lat, int layout, const StepArgs& sa, int taskBase, void* out) {
    if (!out || !groupEligible(lat)) return static_cast<int>(hipErrorInvalidValue);
    GroupEntry e{};
    e.a = makeArgs(lat);
    e.a.tau0 = sa.tau0;
    e.a.lam = sa.magicLambda;
    e.a.pre = sa.smagPrefactor;
    e.a.uIn = sa.uInlet;
    e.a.writeMacro = (sa.writeMacro && lat.rho && lat.u && lat.v && lat.w) ? 1 : 0;
    e.a.stochastic = sa.stochastic;
    e.a.roundSeed = sa.roundSeed;
    e.fromL = layout == kLayoutL ? 1 : 0;
    e.wm = (lat.uwx && lat.uwy && lat.uwz) ? 1 : 0;
    e.taskBase = taskBase;
    std::memcpy(out, &e, sizeof(e));
    return 0;
}

template <class S>
void groupLaunch(dim3 grid, dim3 b, hipStream_t st, const GroupEntry* tab, const GroupLaunch& g, int fromL,
                 int wm) {
    if (fromL < 0 || wm < 0) stepKernelGroupMixed<S><<<grid, b, 0, st>>>(tab, g);
    else if (fromL && wm) stepKernelGroup<S, true, true><<<grid, b, 0, st>>>(tab, g);
    else if (fromL) stepKernelGroup<S, true, false><<<grid, b, 0, st>>>(tab, g);
    else if (wm) stepKernelGroup<S, false, true><<<grid, b, 0, st>>>(tab, g);
    else stepKernelGroup<S, false, false><<<grid, b, 0, st>>>(tab, g);
}

int launchGroupStep(int storage, const void* table, int members, const int* bounds, int wavesPerBlock,
                    int slabOrder, int fromL, int wm, void* stream) {
    if (!table || !bounds || members < 1 || members > kGroupMaxMembers ||
        (wavesPerBlock != 1 && wavesPerBlock != 2 && wavesPerBlock != 4 && wavesPerBlock != 8))
        return static_cast<int>(hipErrorInvalidValue);
    GroupLaunch g{};
    g.members = members;
    for (int m = 0; m <= members; ++m) g.bound[m] = bounds[m];
    g.totalTasks = bounds[members];
    g.wavesPerBlock = wavesPerBlock;
    g.slabOrder = slabOrder ? 1 : 0;
    g.nblocks = (g.totalTasks + wavesPerBlock - 1) / wavesPerBlock;
    g.chunk = (g.nblocks + kXcds - 1) / kXcds;
    if (g.totalTasks <= 0) return 0;
    const unsigned grid = static_cast<unsigned>(g.slabOrder ? g.chunk * kXcds : g.nblocks);
    const dim3 b(static_cast<unsigned>(wavesPerBlock * kWave));
    const hipStream_t st = static_cast<hipStream_t>(stream);
    const auto* tab = static_cast<const GroupEntry*>(table);
    switch (storage) {
        case kStoreF32: groupLaunch<StoreF32>(grid, b, st, tab, g, fromL, wm); break;
        case kStoreF16: groupLaunch<StoreF16>(grid, b, st, tab, g, fromL, wm); break;
        case kStoreBF16: groupLaunch<StoreBF16>(grid, b, st, tab, g, fromL, wm); break;
        case kStoreF16C: groupLaunch<StoreF16C>(grid, b, st, tab, g, fromL, wm); break;
        default: return static_cast<int>(hipErrorInvalidValue);
    }
    return lastError();
}

}  // namespace fc
