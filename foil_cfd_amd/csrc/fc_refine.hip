// Two-level grid refinement on the in-place (AA) population store, gfx950:
// the coarse time-t band snapshot, the coarse-to-fine Interface fill and the
// LDS-tiled fine-to-coarse restriction.
//
// Same coupling as the reference (src/sim/lbm_refine.cu:123-429, sequenced by
// lbm_solver.cpp:1172-1290):
//  * fill: fine Interface cells (or every non-solid fine cell when seeding)
//    get the coarse post-collision populations trilinearly interpolated at
//    their centre (solid corners dropped, weights renormalised), blended
//    between coarse t and t+1 by k/m, split into feq + fneq by the capped
//    moments, fneq projected on the stress space (Latt-Chopard) and rescaled
//    by tau_f / (m tau_c);
//  * restriction: coarse Fluid cells of the box shrunk by 2 get the mean of
//    their m^3 fine children (solid children skipped), the 19-point fneq
//    filter over the neighbours' child averages, the projection, the rescale
//    m tau_c / tau_f, and a blend into the existing coarse value ramping over
//    3 cells from the edge of the restricted box.
// Organised for the in-place store and CDNA4:
//  * the coarse step overwrites time t in place, so the ring of coarse cells
//    the shell fill reads (3 cells wide around the box edge) is copied to a
//    compact FP32 band before the coarse step;
//  * all populations are handled as deviations f - w (interpolation weights
//    sum to one, so interpolating deviations is interpolating f);
//  * the restriction stages each workgroup's child averages (tile plus a
//    one-cell halo) in LDS once, then filters from LDS: 2.5x fine-array reads
//    per coarse cell instead of 19x.
#include <hip/hip_runtime.h>

#include "fc_kcommon.hpp"
#include "fc_refine_math.hpp"

namespace fc {
namespace {

constexpr int kRestrictBlend = 3;  // blend ramp width (lbm_refine.cuh:85)

struct RArgs {
    int x0, y0, x1, y1, m;
    int cnx, cny, cnz;
    long long cn;
    int fnx, fny, fnz;
    long long fn;
    int ew;  // expanded box width (W + 2): columns x0-1 .. x1
    int eh;  // expanded box height (H + 2)
};

RArgs makeRArgs(const RefineGeom& g) {
    RArgs r{};
    r.x0 = g.box.x0;
    r.y0 = g.box.y0;
    r.x1 = g.box.x1;
    r.y1 = g.box.y1;
    r.m = g.factor;
    r.cnx = g.coarse.nx;
    r.cny = g.coarse.ny;
    r.cnz = g.coarse.nz;
    r.cn = g.coarse.cells();
    r.fnx = g.fine.nx;
    r.fny = g.fine.ny;
    r.fnz = g.fine.nz;
    r.fn = g.fine.cells();
    r.ew = g.box.x1 - g.box.x0 + 2;
    r.eh = g.box.y1 - g.box.y0 + 2;
    return r;
}

__device__ inline long long cellOf(int x, int y, int z, int nx, int ny) {
    return static_cast<long long>(x) + static_cast<long long>(nx) * (y + static_cast<long long>(ny) * z);
}

// Post-collision deviation f*_q - w_q of cell (x,y,z) of a level stored in
// `layout` (x / y in range; z periodic).
template <class S>
__device__ inline float readH(const void* fv, long long n, int layout, int q, int x, int y, int z, int nx,
                              int ny, int nz) {
    const typename S::T* f = static_cast<const typename S::T*>(fv);
    if (layout == kLayoutL)
        return static_cast<float>(S::load(f + static_cast<long long>(kOpp[q]) * n + cellOf(x, y, z, nx, ny))) /
               S::K;
    int xx = x + kCx[q], yy = y + kCy[q], zz = z + kCz[q];
    if (zz < 0) zz += nz;
    if (zz >= nz) zz -= nz;
    xx = xx < 0 ? 0 : (xx >= nx ? nx - 1 : xx);  // slots outside x / y are never read back
    yy = yy < 0 ? 0 : (yy >= ny ? ny - 1 : yy);
    return static_cast<float>(S::load(f + static_cast<long long>(q) * n + cellOf(xx, yy, zz, nx, ny))) / S::K;
}

// Store f*_q - w_q of cell (x,y,z) into its slot for `layout`; in layout S a
// slot outside the non-periodic x / y range belongs to no cell and is skipped.
template <class S>
__device__ inline void writeH(void* fv, long long n, int layout, int q, int x, int y, int z, int nx, int ny,
                              int nz, float h) {
    typename S::T* f = static_cast<typename S::T*>(fv);
    using C = typename S::C;
    if (layout == kLayoutL) {
        S::store(f + static_cast<long long>(kOpp[q]) * n + cellOf(x, y, z, nx, ny), C(h) * C(S::K));
        return;
    }
    const int xx = x + kCx[q], yy = y + kCy[q];
    int zz = z + kCz[q];
    if (xx < 0 || xx >= nx || yy < 0 || yy >= ny) return;
    if (zz < 0) zz += nz;
    if (zz >= nz) zz -= nz;
    S::store(f + static_cast<long long>(q) * n + cellOf(xx, yy, zz, nx, ny), C(h) * C(S::K));
}

// Latt-Chopard projection of fneq onto the second-order Hermite space
// (lbm_refine.cu:64-95).
__device__ inline void regularize(float fneq[kQ]) {
    float pxx = 0, pyy = 0, pzz = 0, pxy = 0, pxz = 0, pyz = 0;
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        const float f = fneq[q];
        const float cx = kCx[q], cy = kCy[q], cz = kCz[q];
        pxx += cx * cx * f;
        pyy += cy * cy * f;
        pzz += cz * cz * f;
        pxy += cx * cy * f;
        pxz += cx * cz * f;
        pyz += cy * cz * f;
    }
    const float trace = (1.0f / 3.0f) * (pxx + pyy + pzz);
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        const float cx = kCx[q], cy = kCy[q], cz = kCz[q];
        const float qPi = cx * cx * pxx + cy * cy * pyy + cz * cz * pzz +
                          2.0f * (cx * cy * pxy + cx * cz * pxz + cy * cz * pyz) - trace;
        fneq[q] = 4.5f * wOf<float>(q) * qPi;
    }
}

// ---------------------------------------------------------------------------
// Band snapshot: post-collision deviations of the coarse ring cells at time t.
// band[(q * nz + z) * nband + i] for ring cell i = (bandXY[2i], bandXY[2i+1]).
// ---------------------------------------------------------------------------
template <class S>
__global__ void __launch_bounds__(kBlock) bandSnapshotKernel(const void* f, int layout, RArgs r, const int* bandXY,
                                                            int nband, float* band) {
    const long long t = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x;
    if (t >= static_cast<long long>(nband) * r.cnz) return;
    const int z = static_cast<int>(t / nband);
    const int i = static_cast<int>(t - static_cast<long long>(z) * nband);
    const int x = bandXY[2 * i], y = bandXY[2 * i + 1];
#pragma unroll
    for (int q = 0; q < kQ; ++q)
        band[(static_cast<long long>(q) * r.cnz + z) * nband + i] =
            readH<S>(f, r.cn, layout, q, x, y, z, r.cnx, r.cny, r.cnz);
}

// Fine cell of shell index t (2-cell Interface ring on the x / y faces of every
// fine z plane), or false past the end.
__device__ inline bool shellCell(long long t, const RArgs& r, int& x, int& y, int& z) {
    const long long ring = 4LL * r.fnx + 4LL * (r.fny - 4);
    z = static_cast<int>(t / ring);
    if (z >= r.fnz) return false;
    long long k = t - static_cast<long long>(z) * ring;
    if (k < 2LL * r.fnx) {
        y = static_cast<int>(k / r.fnx);
        x = static_cast<int>(k - static_cast<long long>(y) * r.fnx);
        return true;
    }
    k -= 2LL * r.fnx;
    if (k < 2LL * r.fnx) {
        const int yy = static_cast<int>(k / r.fnx);
        x = static_cast<int>(k - static_cast<long long>(yy) * r.fnx);
        y = r.fny - 2 + yy;
        return true;
    }
    k -= 2LL * r.fnx;
    y = 2 + static_cast<int>(k / 4);
    const int c = static_cast<int>(k % 4);
    x = c < 2 ? c : r.fnx - 4 + c;
    return true;
}

// ---------------------------------------------------------------------------
// Coarse-to-fine fill. FULL: every non-solid fine cell, time levels both the
// live coarse array (seeding). Otherwise: the Interface shell, t from the
// band snapshot and t+1 from the live coarse array.
//
// Which shell populations are ever read: a level's step pulls f*_q of cell x
// only into x + c_q, and only cells inside the shell collide. The band
// snapshot of a nested level reads its parent 3+ cells inside the shell, the
// restriction reads children from fine x, y >= 2 on, and q-LIBB reads only
// links whose target is Solid. So the shell fill stores f*_q(x) only when
// x + c_q lies inside the shell: the 5 inward populations of the inner ring
// (x or y = 1, n - 2), nothing of the outer ring (x or y = 0, n - 1). The
// other shell slots keep older, finite values: a level download returns them
// as they are (compare levels on their Fluid cells), and the NaN watchdog
// skips Interface cells.
// ---------------------------------------------------------------------------
__device__ inline bool insideShell(int x, int y, const RArgs& r) {
    return x >= 2 && y >= 2 && x < r.fnx - 2 && y < r.fny - 2;
}

// Outer ring of the shell: no population of these cells reaches a collided cell.
__device__ inline bool outerRing(int x, int y, const RArgs& r) {
    return x == 0 || y == 0 || x == r.fnx - 1 || y == r.fny - 1;
}

// Moments, fneq projection and rescale of an interpolated shell cell, written
// into the fine slots of layout fLayout (SHELL: only the populations that
// stream into the shell's interior).
template <class S, bool SHELL>
__device__ inline void fillFinish(void* ff, int fLayout, const RArgs& r, int xf, int yf, int zf, const float h[kQ],
                                  float neqScale) {
    const Mom mo = momentsOf(h);
    float fneq[kQ];
#pragma unroll
    for (int q = 0; q < kQ; ++q) fneq[q] = h[q] - heq(q, mo);
    regularize(fneq);
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        if (SHELL && !insideShell(xf + kCx[q], yf + kCy[q], r)) continue;
        writeH<S>(ff, r.fn, fLayout, q, xf, yf, zf, r.fnx, r.fny, r.fnz, heq(q, mo) + neqScale * fneq[q]);
    }
}

template <class S, bool FULL>
__global__ void __launch_bounds__(kBlock) fillKernel(const void* cf, const std::uint8_t* ckind, int cLayout,
                                                    const float* band, const int* bandIndex, int nband,
                                                    void* ff, const std::uint8_t* fkind, int fLayout, RArgs r,
                                                    float timeWeight, float neqScale, float* cache,
                                                    int cacheMode, long long lanes) {
    const long long lb = xcdBlock((lanes + kBlock - 1) / kBlock);
    const long long t = lb * kBlock + threadIdx.x;
    if (t >= lanes) return;
    int xf, yf, zf;
    // Shell interpolation cache: the m sub-steps of one parent step blend the
    // same two interpolated time levels with different weights, so the first
    // stores them (cacheMode 1) and the others reuse them (cacheMode 2).
    // cache[(2q) ns + t] = level t, cache[(2q + 1) ns + t] = level t + 1,
    // cache[38 ns + t] = 1 for a filled shell cell (ns = shell lanes).
    const long long ns = (4LL * r.fnx + 4LL * (r.fny - 4)) * r.fnz;
    float h[kQ];
    if (FULL) {
        if (t >= r.fn) return;
        zf = static_cast<int>(t / (static_cast<long long>(r.fnx) * r.fny));
        const long long rem = t - static_cast<long long>(zf) * r.fnx * r.fny;
        yf = static_cast<int>(rem / r.fnx);
        xf = static_cast<int>(rem - static_cast<long long>(yf) * r.fnx);
        if ((fkind[t] & kKindFlagMask) == kSolid) return;
    } else {
        if (!shellCell(t, r, xf, yf, zf) || outerRing(xf, yf, r)) return;
        if (cacheMode == 2) {
            if (cache[38 * ns + t] == 0.0f) return;
#pragma unroll
            for (int q = 0; q < kQ; ++q) {
                const float s0 = cache[(2 * q) * ns + t], s1 = cache[(2 * q + 1) * ns + t];
                h[q] = s0 + timeWeight * (s1 - s0);
            }
            fillFinish<S, true>(ff, fLayout, r, xf, yf, zf, h, neqScale);
            return;
        }
        if ((fkind[cellOf(xf, yf, zf, r.fnx, r.fny)] & kKindFlagMask) != kInterface) {
            if (cacheMode == 1) cache[38 * ns + t] = 0.0f;
            return;
        }
    }
    const float invM = 1.0f / static_cast<float>(r.m);
    const float gx = static_cast<float>(r.x0) + (static_cast<float>(xf) + 0.5f) * invM - 0.5f;
    const float gy = static_cast<float>(r.y0) + (static_cast<float>(yf) + 0.5f) * invM - 0.5f;
    const float gz = (static_cast<float>(zf) + 0.5f) * invM - 0.5f;
    const int ix = static_cast<int>(floorf(gx)), iy = static_cast<int>(floorf(gy)),
              iz = static_cast<int>(floorf(gz));
    const float tx = gx - ix, ty = gy - iy, tz = gz - iz;
    float wgt[8] = {(1 - tx) * (1 - ty) * (1 - tz), tx * (1 - ty) * (1 - tz), (1 - tx) * ty * (1 - tz),
                    tx * ty * (1 - tz),             (1 - tx) * (1 - ty) * tz, tx * (1 - ty) * tz,
                    (1 - tx) * ty * tz,             tx * ty * tz};
    int cx[8], cy[8], cz[8];
    float wsum = 0.0f;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        cx[c] = min(max(ix + (c & 1), 0), r.cnx - 1);
        cy[c] = min(max(iy + ((c >> 1) & 1), 0), r.cny - 1);
        int zz = iz + ((c >> 2) & 1);
        if (zz < 0) zz += r.cnz;
        if (zz >= r.cnz) zz -= r.cnz;
        cz[c] = zz;
        if ((ckind[cellOf(cx[c], cy[c], cz[c], r.cnx, r.cny)] & kKindFlagMask) == kSolid) wgt[c] = 0.0f;
        wsum += wgt[c];
    }
    if (wsum < 1e-6f) {
        if (!FULL && cacheMode == 1) cache[38 * ns + t] = 0.0f;
        return;
    }
    const float invW = 1.0f / wsum;
    // Ring slot of each corner in the band snapshot (every shell corner lies in
    // the ring; a corner outside it would fall back to the live array).
    int slot[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        wgt[c] *= invW;
        slot[c] = -1;
        if (!FULL) {
            const int ex = cx[c] - (r.x0 - 1), ey = cy[c] - (r.y0 - 1);
            if (ex >= 0 && ex < r.ew && ey >= 0 && ey < r.eh) slot[c] = bandIndex[ey * r.ew + ex];
        }
    }
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        float s0 = 0.0f, s1 = 0.0f;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            if (wgt[c] == 0.0f) continue;  // solid corner: skip its (unused) slots
            const float v1 = readH<S>(cf, r.cn, cLayout, q, cx[c], cy[c], cz[c], r.cnx, r.cny, r.cnz);
            const float v0 =
                slot[c] < 0 ? v1 : band[(static_cast<long long>(q) * r.cnz + cz[c]) * nband + slot[c]];
            s0 += wgt[c] * v0;
            s1 += wgt[c] * v1;
        }
        if (!FULL && cacheMode == 1) {
            cache[(2 * q) * ns + t] = s0;
            cache[(2 * q + 1) * ns + t] = s1;
        }
        h[q] = s0 + timeWeight * (s1 - s0);
    }
    if (!FULL && cacheMode == 1) cache[38 * ns + t] = 1.0f;
    fillFinish<S, !FULL>(ff, fLayout, r, xf, yf, zf, h, neqScale);
}

// Shell interpolation for factor 2, grouped by shared corners. The fine
// cells 2k - 1 and 2k of any axis have the same floor((i + 1/2) / 2 - 1/2),
// so the shell splits into groups of up to four cells with one set of eight
// coarse corners: an x pair of a shell row, or a y pair of a shell column,
// times a z pair (2k - 1, 2k), z = -1 being plane fnz - 1. One lane per
// group loads each corner value once for all members and writes both
// interpolated time levels of every member into the fill cache, with each
// member's own weights and the per-cell summation order of fillKernel, so
// the cache holds the same bits fillKernel's first sub-step stores; the
// cached finish (cacheMode 2) then writes the fine populations.
// Groups g = (kz * 4 + lane kind) ...: per z pair, 4 rows of w + 1 x pairs
// (x = 2kx - 1 + d), then 4 side columns of h - 1 y pairs (y = 2ky - 1 + d,
// ky = 1 .. h - 1, members kept when 2 <= y < fny - 2). The launch has
// kInterpRows grid rows, each doing the gathers of a share of the 19
// populations.
constexpr unsigned kInterpRows = 4;

template <class S>
__global__ void __launch_bounds__(kBlock) shellInterp2Kernel(const void* cf, const std::uint8_t* ckind, int cLayout,
                                                            const float* band, const int* bandIndex, int nband,
                                                            const std::uint8_t* fkind, RArgs r, float* cache,
                                                            long long groups) {
    const long long g = xcdBlock((groups + kBlock - 1) / kBlock) * kBlock + threadIdx.x;
    if (g >= groups) return;
    const int w = r.x1 - r.x0, h = r.y1 - r.y0;
    const int perZ = 4 * (w + h);
    const int kz = static_cast<int>(g / perZ);
    const int rem = static_cast<int>(g - static_cast<long long>(kz) * perZ);
    const long long ring = 4LL * r.fnx + 4LL * (r.fny - 4);
    const long long ns = ring * r.fnz;
    // members (d, e): pair member d along x (rows) or y (columns), z member e
    int mx[2], my[2];
    bool ok[2];
    long long off[2];  // shell offset within a z plane
    if (rem < 4 * (w + 1)) {
        const int row = rem / (w + 1), kx = rem - row * (w + 1);
        const int y = row < 2 ? row : r.fny - 4 + row;
        for (int d = 0; d < 2; ++d) {
            mx[d] = 2 * kx - 1 + d;
            my[d] = y;
            ok[d] = mx[d] >= 0 && mx[d] < r.fnx && !outerRing(mx[d], y, r);
            off[d] = (row < 2 ? static_cast<long long>(row) * r.fnx
                              : 2LL * r.fnx + static_cast<long long>(row - 2) * r.fnx) + mx[d];
        }
    } else {
        const int rem2 = rem - 4 * (w + 1);
        const int side = rem2 / (h - 1), ky = 1 + rem2 - side * (h - 1);
        const int x = side < 2 ? side : r.fnx - 4 + side;
        for (int d = 0; d < 2; ++d) {
            mx[d] = x;
            my[d] = 2 * ky - 1 + d;
            ok[d] = my[d] >= 2 && my[d] < r.fny - 2 && !outerRing(x, my[d], r);
            off[d] = 4LL * r.fnx + 4LL * (my[d] - 2) + side;
        }
    }
    const int mz[2] = {kz == 0 ? r.fnz - 1 : 2 * kz - 1, 2 * kz};
    // shared corners, from the representative cell (2 kx | x, 2 ky | y, 2 kz)
    const float invM = 1.0f / static_cast<float>(r.m);
    const int rx = ok[1] ? mx[1] : mx[0], ry = ok[1] ? my[1] : my[0];
    const int ix = static_cast<int>(floorf(static_cast<float>(r.x0) + (static_cast<float>(rx) + 0.5f) * invM - 0.5f));
    const int iy = static_cast<int>(floorf(static_cast<float>(r.y0) + (static_cast<float>(ry) + 0.5f) * invM - 0.5f));
    const int iz = kz - 1;
    int cx[8], cy[8], cz[8], slot[8];
    bool solid[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        cx[c] = min(max(ix + (c & 1), 0), r.cnx - 1);
        cy[c] = min(max(iy + ((c >> 1) & 1), 0), r.cny - 1);
        int zz = iz + ((c >> 2) & 1);
        if (zz < 0) zz += r.cnz;
        if (zz >= r.cnz) zz -= r.cnz;
        cz[c] = zz;
        solid[c] = (ckind[cellOf(cx[c], cy[c], cz[c], r.cnx, r.cny)] & kKindFlagMask) == kSolid;
        slot[c] = -1;
        const int ex = cx[c] - (r.x0 - 1), ey = cy[c] - (r.y0 - 1);
        if (ex >= 0 && ex < r.ew && ey >= 0 && ey < r.eh) slot[c] = bandIndex[ey * r.ew + ex];
    }
    // per member: shell index and normalised weights (as fillKernel), or t = -1
    long long t[4];
    float wgt[4][8];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const int d = m & 1, e = m >> 1;
        t[m] = -1;
        if (!ok[d]) continue;
        const int xf = mx[d], yf = my[d], zf = mz[e];
        const long long tm = static_cast<long long>(zf) * ring + off[d];
        if ((fkind[cellOf(xf, yf, zf, r.fnx, r.fny)] & kKindFlagMask) != kInterface) {
            if (blockIdx.y == 0) cache[38 * ns + tm] = 0.0f;
            continue;
        }
        const float gx = static_cast<float>(r.x0) + (static_cast<float>(xf) + 0.5f) * invM - 0.5f;
        const float gy = static_cast<float>(r.y0) + (static_cast<float>(yf) + 0.5f) * invM - 0.5f;
        const float gz = (static_cast<float>(zf) + 0.5f) * invM - 0.5f;
        const int jx = static_cast<int>(floorf(gx)), jy = static_cast<int>(floorf(gy)),
                  jz = static_cast<int>(floorf(gz));
        const float tx = gx - jx, ty = gy - jy, tz = gz - jz;
        float wv[8] = {(1 - tx) * (1 - ty) * (1 - tz), tx * (1 - ty) * (1 - tz), (1 - tx) * ty * (1 - tz),
                       tx * ty * (1 - tz),             (1 - tx) * (1 - ty) * tz, tx * (1 - ty) * tz,
                       (1 - tx) * ty * tz,             tx * ty * tz};
        float wsum = 0.0f;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            if (solid[c]) wv[c] = 0.0f;
            wsum += wv[c];
        }
        if (wsum < 1e-6f) {
            if (blockIdx.y == 0) cache[38 * ns + tm] = 0.0f;
            continue;
        }
        const float invW = 1.0f / wsum;
#pragma unroll
        for (int c = 0; c < 8; ++c) wgt[m][c] = wv[c] * invW;
        t[m] = tm;
    }
    if (t[0] < 0 && t[1] < 0 && t[2] < 0 && t[3] < 0) return;
    // grid row y takes the populations [q0, q1): the groups' set-up repeats
    // per row, the gathers and sums split, so more waves hide the latency
    const int q0 = static_cast<int>(kQ * blockIdx.y / gridDim.y);
    const int q1 = static_cast<int>(kQ * (blockIdx.y + 1) / gridDim.y);
    for (int q = q0; q < q1; ++q) {
        float v0[8], v1[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            v0[c] = v1[c] = 0.0f;
            if (solid[c]) continue;
            v1[c] = readH<S>(cf, r.cn, cLayout, q, cx[c], cy[c], cz[c], r.cnx, r.cny, r.cnz);
            v0[c] = slot[c] < 0 ? v1[c] : band[(static_cast<long long>(q) * r.cnz + cz[c]) * nband + slot[c]];
        }
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            if (t[m] < 0) continue;
            float s0 = 0.0f, s1 = 0.0f;
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                if (wgt[m][c] == 0.0f) continue;
                s0 += wgt[m][c] * v0[c];
                s1 += wgt[m][c] * v1[c];
            }
            cache[(2 * q) * ns + t[m]] = s0;
            cache[(2 * q + 1) * ns + t[m]] = s1;
        }
    }
    if (blockIdx.y == 0) {
#pragma unroll
        for (int m = 0; m < 4; ++m)
            if (t[m] >= 0) cache[38 * ns + t[m]] = 1.0f;
    }
}

// ---------------------------------------------------------------------------
// Fine-to-coarse restriction. With a scratch array (the solver's path) a
// child-average pass writes one record per coarse cell of the restricted box
// and an LDS-tiled filter pass reads them back; without one, a single LDS-
// tiled kernel averages the children of its tile plus halo itself. The
// staged filter tile (TX, TY, TZ) is FC_RESTRICT_TILE.
// ---------------------------------------------------------------------------
constexpr int kTX = 16, kTY = 4, kTZ = 4;
constexpr int kHX = kTX + 2, kHY = kTY + 2, kHZ = kTZ + 2;
constexpr int kHalo = kHX * kHY * kHZ;  // 648 cells, 19 floats each = 49 KB
#ifndef FC_RESTRICT_TILE
#define FC_RESTRICT_TILE 16, 8, 4
#endif
constexpr int kStagedTile[3] = {FC_RESTRICT_TILE};
constexpr int kStagedTX = kStagedTile[0], kStagedTY = kStagedTile[1], kStagedTZ = kStagedTile[2];

// Child-average pass of the staged restriction: one wavefront per row
// segment of 64 / m coarse cells of the restricted box, m lanes per coarse
// cell, lane l on fine column m (xc - x0) + l % m, so every population load
// of the wave is one contiguous run of fine x. Each lane sums its m x m
// children in (y, z), the m lanes of a cell are combined by cross-lane reads,
// and the cell's first lane stores its record (fneq = h_avg - heq(h_avg), the
// valid mark -- 1 when the cell has a non-solid child -- and the moments rho,
// ux, uy, uz) at ri = (xc - rx0) + rw ((yc - ry0) + rh zc) in the full or the
// compact format of the storage (storeRecord).
template <class S>
__global__ void __launch_bounds__(kBlock) childFneqKernel(const void* ff, const std::uint8_t* fkind, int fLayout,
                                                         RArgs r, float* scratch) {
    const int rx0 = r.x0 + kRestrictBand, ry0 = r.y0 + kRestrictBand;
    const int rw = r.x1 - kRestrictBand - rx0, rh = r.y1 - kRestrictBand - ry0;
    const int m = r.m;
    const int cpw = kWave / m;  // coarse cells per wave
    const int nseg = (rw + cpw - 1) / cpw;
    const long long wave = static_cast<long long>(blockIdx.x) * kWavesPerBlock + (threadIdx.x >> 6);
    if (wave >= static_cast<long long>(nseg) * rh * r.cnz) return;  // wave-uniform
    const int lane = static_cast<int>(threadIdx.x) & (kWave - 1);
    const int seg = static_cast<int>(wave % nseg);
    const long long row = wave / nseg;
    const int yc = ry0 + static_cast<int>(row % rh);
    const int zc = static_cast<int>(row / rh);
    const int k = lane / m, lx = lane - k * m;
    const int xc = rx0 + seg * cpw + k;
    const bool live = k < cpw && xc < rx0 + rw;
    float h[kQ];
#pragma unroll
    for (int q = 0; q < kQ; ++q) h[q] = 0.0f;
    int nf = 0;
    if (live) {
        const int fx = m * (xc - r.x0) + lx;
        for (int kk = 0; kk < m; ++kk)
            for (int j = 0; j < m; ++j) {
                const int fy = m * (yc - r.y0) + j, fz = m * zc + kk;
                if ((fkind[cellOf(fx, fy, fz, r.fnx, r.fny)] & kKindFlagMask) == kSolid) continue;
                ++nf;
#pragma unroll
                for (int q = 0; q < kQ; ++q) h[q] += readH<S>(ff, r.fn, fLayout, q, fx, fy, fz, r.fnx, r.fny, r.fnz);
            }
    }
    // the m lanes of a coarse cell: the first gathers the others' sums
    for (int i = 1; i < m; ++i) {
        const int src = min(lane + i, kWave - 1);
#pragma unroll
        for (int q = 0; q < kQ; ++q) {
            const float o = __shfl(h[q], src, kWave);
            if (lx == 0) h[q] += o;
        }
        const int on = __shfl(nf, src, kWave);
        if (lx == 0) nf += on;
    }
    if (!live || lx != 0) return;
    const long long nR = static_cast<long long>(rw) * rh * r.cnz;
    const long long ri = (xc - rx0) + static_cast<long long>(rw) * ((yc - ry0) + static_cast<long long>(rh) * zc);
    float rec[4 * kRecordPlanes];
    childRecord(h, nf, rec);
    storeRecord<RecordFormat<S>::kCompact>(reinterpret_cast<float4*>(scratch), nR, ri, rec);
}

// Restriction of one coarse cell from its filtered fneq (already divided by
// the number of valid cells in the 19-point neighbourhood) and the child-
// averaged moments: Latt-Chopard projection, rescale to the coarse tau, edge
// blend into the existing populations, optional macro store.
template <class S>
__device__ inline void restrictFinish(void* cf, int cLayout, const RArgs& r, int xc, int yc, int zc, long long cc,
                                      int dEdge, float* fneq, const Mom& mo, float neqScaleInv, float* mRho,
                                      float* mU, float* mV, float* mW) {
    const float blend = fminf(1.0f, static_cast<float>(dEdge + 1) / static_cast<float>(kRestrictBlend + 1));
    regularize(fneq);
    float br = 0, bjx = 0, bjy = 0, bjz = 0;
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        const float existing = readH<S>(cf, r.cn, cLayout, q, xc, yc, zc, r.cnx, r.cny, r.cnz);
        const float restricted = heq(q, mo) + neqScaleInv * fneq[q];
        const float blended = existing + blend * (restricted - existing);
        writeH<S>(cf, r.cn, cLayout, q, xc, yc, zc, r.cnx, r.cny, r.cnz, blended);
        br += blended;
        bjx += kCx[q] * blended;
        bjy += kCy[q] * blended;
        bjz += kCz[q] * blended;
    }
    if (mRho) {
        const float rr = fmaxf(1.0f + br, 0.05f);
        const float inv = 1.0f / rr;
        mRho[cc] = rr;
        mU[cc] = bjx * inv;
        mV[cc] = bjy * inv;
        mW[cc] = bjz * inv;
    }
}

// Staged restriction filter over the planes of the child-average pass. A
// workgroup covers TX x TY x TZ coarse cells of the restricted box and
// stages the tile plus a one-cell halo through a double-buffered LDS tile,
// one plane (four populations) at a time, the valid plane first; while a
// plane is summed the next one is in flight. Each thread accumulates the
// 19-point filter of its own cell in the neighbour order of the unstaged
// kernel, so both give identical bits.
template <class S, int TX, int TY, int TZ>
__global__ void __launch_bounds__(TX * TY * TZ) restrictStagedKernel(void* cf, const std::uint8_t* ckind,
                                                                    int cLayout, RArgs r, float neqScaleInv,
                                                                    float* mRho, float* mU, float* mV, float* mW,
                                                                    const float4* scratch4) {
    constexpr int kT = TX * TY * TZ;
    constexpr int HX = TX + 2, HY = TY + 2, HZ = TZ + 2, kH = HX * HY * HZ;
    constexpr int kPer = (kH + kT - 1) / kT;
    __shared__ float4 sRec[2][kH];
    const int rx0 = r.x0 + kRestrictBand, ry0 = r.y0 + kRestrictBand;
    const int rx1 = r.x1 - kRestrictBand, ry1 = r.y1 - kRestrictBand;
    const int rw = rx1 - rx0, rh = ry1 - ry0;
    const long long nR = static_cast<long long>(rw) * rh * r.cnz;
    const int ntx = (rw + TX - 1) / TX, nty = (rh + TY - 1) / TY;
    const long long nb = static_cast<long long>(ntx) * nty * ((r.cnz + TZ - 1) / TZ);
    const long long lbk = xcdBlock(nb);
    if (lbk >= nb) return;  // workgroup-uniform
    const int b = static_cast<int>(lbk);
    const int bx = b % ntx, by = (b / ntx) % nty, bz = b / (ntx * nty);
    const int ox = rx0 + bx * TX, oy = ry0 + by * TY, oz = bz * TZ;
    const int t = static_cast<int>(threadIdx.x);

    // record index of each halo slot this thread stages (-1: outside the box)
    long long src[kPer];
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const int i = t + j * kT;
        src[j] = -1;
        if (i < kH) {
            const int hx = i % HX, hy = (i / HX) % HY, hz = i / (HX * HY);
            const int xc = ox + hx - 1, yc = oy + hy - 1;
            const int zc = ((oz + hz - 1) % r.cnz + r.cnz) % r.cnz;
            if (xc >= rx0 && xc < rx1 && yc >= ry0 && yc < ry1)
                src[j] = (xc - rx0) + static_cast<long long>(rw) * ((yc - ry0) + static_cast<long long>(rh) * zc);
        }
    }
    float4 pre[kPer];
    auto fetch = [&](int c) {
#pragma unroll
        for (int j = 0; j < kPer; ++j)
            pre[j] = src[j] >= 0 ? scratch4[c * nR + src[j]] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    };
    auto stage = [&](int buf) {
#pragma unroll
        for (int j = 0; j < kPer; ++j)
            if (t + j * kT < kH) sRec[buf][t + j * kT] = pre[j];
    };

    const int lx = t % TX, ly = (t / TX) % TY, lz = t / (TX * TY);
    const int xc = ox + lx, yc = oy + ly, zc = oz + lz;
    const int hc = (lx + 1) + HX * ((ly + 1) + HY * (lz + 1));
    const bool inBox = xc < rx1 && yc < ry1 && zc < r.cnz;
    const long long cc = inBox ? cellOf(xc, yc, zc, r.cnx, r.cny) : 0;
    const long long own = inBox ? (xc - rx0) + static_cast<long long>(rw) * ((yc - ry0) + static_cast<long long>(rh) * zc) : 0;
    bool active = inBox && (ckind[cc] & kKindFlagMask) == kFluid;
    const float4 mom = active ? scratch4[5 * nR + own] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);

    fetch(4);
    stage(0);
    __syncthreads();
    unsigned nmask = 0;
    if (active) {
#pragma unroll
        for (int n = 0; n < kQ; ++n)
            if (sRec[0][hc + kCx[n] + HX * (kCy[n] + HY * kCz[n])].w != 0.0f) nmask |= 1u << n;
        active = (nmask & 1u) != 0;
    }
    float fneq[kQ];
    constexpr int kOrder[5] = {4, 0, 1, 2, 3};
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        if (k < 4) fetch(kOrder[k + 1]);
        if (active) {
            const float4* s = sRec[k & 1];
            float4 a = s[hc];
#pragma unroll
            for (int n = 1; n < kQ; ++n) {
                if (!(nmask & (1u << n))) continue;
                const float4 v = s[hc + kCx[n] + HX * (kCy[n] + HY * kCz[n])];
                a.x += v.x;
                a.y += v.y;
                a.z += v.z;
                a.w += v.w;
            }
            const int q0 = 4 * kOrder[k];
            fneq[q0] = a.x;
            fneq[q0 + 1] = a.y;
            fneq[q0 + 2] = a.z;
            if (q0 + 3 < kQ) fneq[q0 + 3] = a.w;
        }
        if (k < 4) {
            stage((k + 1) & 1);
            __syncthreads();
        }
    }
    if (!active) return;
    const float invS = 1.0f / static_cast<float>(__popc(nmask));
#pragma unroll
    for (int q = 0; q < kQ; ++q) fneq[q] *= invS;
    const int dEdge = min(min(xc - rx0, rx1 - 1 - xc), min(yc - ry0, ry1 - 1 - yc));
    restrictFinish<S>(cf, cLayout, r, xc, yc, zc, cc, dEdge, fneq, Mom{mom.x, mom.y, mom.z, mom.w}, neqScaleInv,
                      mRho, mU, mV, mW);
}

// The staged filter over compact records (16-bit storages): the three
// planes of eight halves are fetched into registers one at a time (the next
// one in flight while the current is summed) and each is decoded once into
// both LDS buffers, four values per buffer; the plane holding the valid mark
// (fneq[16..18] + valid) first. Per value the neighbour order and the
// arithmetic are those of the full-format filter, on the records' rounded
// values; the sums are unscaled once at the end (an exact power of two).
template <class S, int TX, int TY, int TZ>
__global__ void __launch_bounds__(TX * TY * TZ) restrictStagedCompactKernel(void* cf, const std::uint8_t* ckind,
                                                                           int cLayout, RArgs r, float neqScaleInv,
                                                                           float* mRho, float* mU, float* mV,
                                                                           float* mW, const float4* scratch4) {
    constexpr int kT = TX * TY * TZ;
    constexpr int HX = TX + 2, HY = TY + 2, HZ = TZ + 2, kH = HX * HY * HZ;
    constexpr int kPer = (kH + kT - 1) / kT;
    __shared__ float4 sRec[2][kH];
    const int rx0 = r.x0 + kRestrictBand, ry0 = r.y0 + kRestrictBand;
    const int rx1 = r.x1 - kRestrictBand, ry1 = r.y1 - kRestrictBand;
    const int rw = rx1 - rx0, rh = ry1 - ry0;
    const long long nR = static_cast<long long>(rw) * rh * r.cnz;
    const int ntx = (rw + TX - 1) / TX, nty = (rh + TY - 1) / TY;
    const long long nb = static_cast<long long>(ntx) * nty * ((r.cnz + TZ - 1) / TZ);
    const long long lbk = xcdBlock(nb);
    if (lbk >= nb) return;  // workgroup-uniform
    const int b = static_cast<int>(lbk);
    const int bx = b % ntx, by = (b / ntx) % nty, bz = b / (ntx * nty);
    const int ox = rx0 + bx * TX, oy = ry0 + by * TY, oz = bz * TZ;
    const int t = static_cast<int>(threadIdx.x);

    // record index of each halo slot this thread stages (-1: outside the box)
    long long src[kPer];
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const int i = t + j * kT;
        src[j] = -1;
        if (i < kH) {
            const int hx = i % HX, hy = (i / HX) % HY, hz = i / (HX * HY);
            const int xc = ox + hx - 1, yc = oy + hy - 1;
            const int zc = ((oz + hz - 1) % r.cnz + r.cnz) % r.cnz;
            if (xc >= rx0 && xc < rx1 && yc >= ry0 && yc < ry1)
                src[j] = (xc - rx0) + static_cast<long long>(rw) * ((yc - ry0) + static_cast<long long>(rh) * zc);
        }
    }
    const uint4* const h4 = reinterpret_cast<const uint4*>(scratch4 + nR);
    uint4 pre[kPer];
    auto fetch = [&](int c) {
#pragma unroll
        for (int j = 0; j < kPer; ++j) pre[j] = src[j] >= 0 ? h4[c * nR + src[j]] : make_uint4(0u, 0u, 0u, 0u);
    };
    auto stage = [&]() {
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            if (t + j * kT >= kH) continue;
            const uint4 v = pre[j];
            sRec[0][t + j * kT] = make_float4(recUnpack(v, 0), recUnpack(v, 1), recUnpack(v, 2), recUnpack(v, 3));
            sRec[1][t + j * kT] = make_float4(recUnpack(v, 4), recUnpack(v, 5), recUnpack(v, 6), recUnpack(v, 7));
        }
    };

    const int lx = t % TX, ly = (t / TX) % TY, lz = t / (TX * TY);
    const int xc = ox + lx, yc = oy + ly, zc = oz + lz;
    const int hc = (lx + 1) + HX * ((ly + 1) + HY * (lz + 1));
    const bool inBox = xc < rx1 && yc < ry1 && zc < r.cnz;
    const long long cc = inBox ? cellOf(xc, yc, zc, r.cnx, r.cny) : 0;
    const long long own = inBox ? (xc - rx0) + static_cast<long long>(rw) * ((yc - ry0) + static_cast<long long>(rh) * zc) : 0;
    bool active = inBox && (ckind[cc] & kKindFlagMask) == kFluid;
    const float4 mom = active ? scratch4[own] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);

    fetch(2);
    stage();
    __syncthreads();
    unsigned nmask = 0;
    if (active) {
#pragma unroll
        for (int n = 0; n < kQ; ++n)
            if (sRec[0][hc + kCx[n] + HX * (kCy[n] + HY * kCz[n])].w != 0.0f) nmask |= 1u << n;
        active = (nmask & 1u) != 0;
    }
    float fneq[kQ];
    constexpr int kOrder[3] = {2, 0, 1};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (k < 2) fetch(kOrder[k + 1]);
        if (active) {
#pragma unroll
            for (int hs = 0; hs < 2; ++hs) {
                if (kOrder[k] == 2 && hs == 1) continue;  // halves 4..7 of the valid plane are padding
                const float4* sp = sRec[hs];
                float4 a = sp[hc];
#pragma unroll
                for (int n = 1; n < kQ; ++n) {
                    if (!(nmask & (1u << n))) continue;
                    const float4 v = sp[hc + kCx[n] + HX * (kCy[n] + HY * kCz[n])];
                    a.x += v.x;
                    a.y += v.y;
                    a.z += v.z;
                    a.w += v.w;
                }
                const int q0 = 8 * kOrder[k] + 4 * hs;
                fneq[q0] = a.x;
                fneq[q0 + 1] = a.y;
                fneq[q0 + 2] = a.z;
                if (q0 + 3 < kQ) fneq[q0 + 3] = a.w;
            }
        }
        if (k < 2) {
            __syncthreads();
            stage();
            __syncthreads();
        }
    }
    if (!active) return;
    const float invS = 1.0f / static_cast<float>(__popc(nmask));
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        fneq[q] *= invS;
        fneq[q] *= kRecInvScale;
    }
    const int dEdge = min(min(xc - rx0, rx1 - 1 - xc), min(yc - ry0, ry1 - 1 - yc));
    restrictFinish<S>(cf, cLayout, r, xc, yc, zc, cc, dEdge, fneq, Mom{mom.x, mom.y, mom.z, mom.w}, neqScaleInv,
                      mRho, mU, mV, mW);
}

// Unstaged restriction (no scratch): a workgroup covers kTX x kTY x kTZ
// coarse cells; phase 1 averages the children of the tile plus a one-cell
// halo into LDS (fneq, and the moments of the tile's own cells), phase 2
// filters, projects, rescales and blends one coarse cell per thread.
template <class S>
__global__ void __launch_bounds__(kBlock) restrictDirectKernel(const void* ff, const std::uint8_t* fkind,
                                                              int fLayout, void* cf, const std::uint8_t* ckind,
                                                              int cLayout, RArgs r, float neqScaleInv, float* mRho,
                                                              float* mU, float* mV, float* mW) {
    __shared__ float sFneq[kQ][kHalo];
    __shared__ unsigned char sValid[kHalo];
    __shared__ float sMom[4][kTX * kTY * kTZ];
    const int rx0 = r.x0 + kRestrictBand, ry0 = r.y0 + kRestrictBand;
    const int rx1 = r.x1 - kRestrictBand, ry1 = r.y1 - kRestrictBand;
    const int ntx = (rx1 - rx0 + kTX - 1) / kTX, nty = (ry1 - ry0 + kTY - 1) / kTY;
    const int b = blockIdx.x;
    const int bx = b % ntx, by = (b / ntx) % nty, bz = b / (ntx * nty);
    const int ox = rx0 + bx * kTX, oy = ry0 + by * kTY, oz = bz * kTZ;

    for (int i = threadIdx.x; i < kHalo; i += kBlock) {
        const int hx = i % kHX, hy = (i / kHX) % kHY, hz = i / (kHX * kHY);
        const int xc = ox + hx - 1, yc = oy + hy - 1;
        int zc = oz + hz - 1;
        zc = (zc % r.cnz + r.cnz) % r.cnz;
        unsigned char ok = 0;
        if (xc >= rx0 && xc < rx1 && yc >= ry0 && yc < ry1) {
            float h[kQ];
#pragma unroll
            for (int q = 0; q < kQ; ++q) h[q] = 0.0f;
            int nf = 0;
            const int fx = r.m * (xc - r.x0), fy = r.m * (yc - r.y0), fz = r.m * zc;
            for (int k = 0; k < r.m; ++k)
                for (int j = 0; j < r.m; ++j)
                    for (int l = 0; l < r.m; ++l) {
                        const int x = fx + l, y = fy + j, z = fz + k;
                        if ((fkind[cellOf(x, y, z, r.fnx, r.fny)] & kKindFlagMask) == kSolid) continue;
                        ++nf;
#pragma unroll
                        for (int q = 0; q < kQ; ++q) h[q] += readH<S>(ff, r.fn, fLayout, q, x, y, z, r.fnx, r.fny, r.fnz);
                    }
            if (nf > 0) {
                ok = 1;
                const float invN = 1.0f / static_cast<float>(nf);
#pragma unroll
                for (int q = 0; q < kQ; ++q) h[q] *= invN;
                const Mom mo = momentsOf(h);
#pragma unroll
                for (int q = 0; q < kQ; ++q) sFneq[q][i] = h[q] - heq(q, mo);
                if (hx >= 1 && hx <= kTX && hy >= 1 && hy <= kTY && hz >= 1 && hz <= kTZ) {
                    const int o = (hx - 1) + kTX * ((hy - 1) + kTY * (hz - 1));
                    sMom[0][o] = mo.rho;
                    sMom[1][o] = mo.ux;
                    sMom[2][o] = mo.uy;
                    sMom[3][o] = mo.uz;
                }
            }
        }
        sValid[i] = ok;
    }
    __syncthreads();

    const int lx = threadIdx.x % kTX, ly = (threadIdx.x / kTX) % kTY, lz = threadIdx.x / (kTX * kTY);
    const int xc = ox + lx, yc = oy + ly, zc = oz + lz;
    if (xc >= rx1 || yc >= ry1 || zc >= r.cnz) return;
    const long long cc = cellOf(xc, yc, zc, r.cnx, r.cny);
    if ((ckind[cc] & kKindFlagMask) != kFluid) return;
    const int hc = (lx + 1) + kHX * ((ly + 1) + kHY * (lz + 1));
    if (!sValid[hc]) return;
    float fneq[kQ];
#pragma unroll
    for (int q = 0; q < kQ; ++q) fneq[q] = sFneq[q][hc];
    int nS = 1;
    for (int n = 1; n < kQ; ++n) {
        const int hn = hc + kCx[n] + kHX * (kCy[n] + kHY * kCz[n]);
        if (!sValid[hn]) continue;
#pragma unroll
        for (int q = 0; q < kQ; ++q) fneq[q] += sFneq[q][hn];
        ++nS;
    }
    const float invS = 1.0f / static_cast<float>(nS);
#pragma unroll
    for (int q = 0; q < kQ; ++q) fneq[q] *= invS;
    const int o = lx + kTX * (ly + kTY * lz);
    const int dEdge = min(min(xc - rx0, rx1 - 1 - xc), min(yc - ry0, ry1 - 1 - yc));
    restrictFinish<S>(cf, cLayout, r, xc, yc, zc, cc, dEdge, fneq, Mom{sMom[0][o], sMom[1][o], sMom[2][o], sMom[3][o]},
                      neqScaleInv, mRho, mU, mV, mW);
}

template <class S>
struct RefineOps {
    static int snapshot(const LatticeView& c, int layout, const RArgs& r, const int* bandXY, int nband,
                        float* band, hipStream_t st) {
        const long long n = static_cast<long long>(nband) * r.cnz;
        bandSnapshotKernel<S><<<static_cast<unsigned>((n + kBlock - 1) / kBlock), kBlock, 0, st>>>(
            c.f, layout, r, bandXY, nband, band);
        return static_cast<int>(hipGetLastError());
    }
    static int fill(const LatticeView& c, int cLayout, const float* band, const int* bandIndex, int nband,
                    const LatticeView& f, int fLayout, const RArgs& r, float tw, float neqScale, bool full,
                    float* cache, int cacheMode, hipStream_t st) {
        if (full) {
            fillKernel<S, true><<<xcdGrid((r.fn + kBlock - 1) / kBlock), kBlock, 0, st>>>(
                c.f, c.kind, cLayout, band, bandIndex, nband, f.f, f.kind, fLayout, r, tw, neqScale, nullptr, 0, r.fn);
        } else {
            const long long n = (4LL * r.fnx + 4LL * (r.fny - 4)) * r.fnz;
            if (cache && cacheMode == 1 && r.m == 2) {
                // grouped interpolation into the cache, then the cached finish
                const long long groups = 4LL * ((r.x1 - r.x0) + (r.y1 - r.y0)) * r.cnz;
                const dim3 grid(xcdGrid((groups + kBlock - 1) / kBlock), kInterpRows);
                shellInterp2Kernel<S><<<grid, kBlock, 0, st>>>(
                    c.f, c.kind, cLayout, band, bandIndex, nband, f.kind, r, cache, groups);
                cacheMode = 2;
            }
            fillKernel<S, false><<<xcdGrid((n + kBlock - 1) / kBlock), kBlock, 0, st>>>(
                c.f, c.kind, cLayout, band, bandIndex, nband, f.f, f.kind, fLayout, r, tw, neqScale,
                cache ? cache : nullptr, cache ? cacheMode : 0, n);
        }
        return static_cast<int>(hipGetLastError());
    }
    static int restrictTo(const LatticeView& f, int fLayout, const LatticeView& c, int cLayout, const RArgs& r,
                          float neqScaleInv, float* rho, float* u, float* v, float* w, float* scratch,
                          bool recordsReady, hipStream_t st) {
        const int rw = r.x1 - r.x0 - 2 * kRestrictBand, rh = r.y1 - r.y0 - 2 * kRestrictBand;
        if (rw <= 0 || rh <= 0) return static_cast<int>(hipErrorInvalidValue);
        const bool macro = rho && u && v && w;
        if (scratch) {
            const int cpw = kWave / r.m;
            const long long waves = static_cast<long long>((rw + cpw - 1) / cpw) * rh * r.cnz;
            if (!recordsReady)
                childFneqKernel<S><<<static_cast<unsigned>((waves + kWavesPerBlock - 1) / kWavesPerBlock), kBlock,
                                     0, st>>>(f.f, f.kind, fLayout, r, scratch);
            constexpr int TX = kStagedTX, TY = kStagedTY, TZ = kStagedTZ;
            const long long blocks = static_cast<long long>((rw + TX - 1) / TX) * ((rh + TY - 1) / TY) *
                                     ((r.cnz + TZ - 1) / TZ);
            if constexpr (RecordFormat<S>::kCompact)
                restrictStagedCompactKernel<S, TX, TY, TZ><<<xcdGrid(blocks), TX * TY * TZ, 0, st>>>(
                    c.f, c.kind, cLayout, r, neqScaleInv, macro ? rho : nullptr, u, v, w,
                    reinterpret_cast<const float4*>(scratch));
            else
                restrictStagedKernel<S, TX, TY, TZ><<<xcdGrid(blocks), TX * TY * TZ, 0, st>>>(
                    c.f, c.kind, cLayout, r, neqScaleInv, macro ? rho : nullptr, u, v, w,
                    reinterpret_cast<const float4*>(scratch));
        } else {
            const unsigned blocks = static_cast<unsigned>(((rw + kTX - 1) / kTX) * ((rh + kTY - 1) / kTY) *
                                                          ((r.cnz + kTZ - 1) / kTZ));
            restrictDirectKernel<S><<<blocks, kBlock, 0, st>>>(f.f, f.kind, fLayout, c.f, c.kind, cLayout, r,
                                                               neqScaleInv, macro ? rho : nullptr, u, v, w);
        }
        return static_cast<int>(hipGetLastError());
    }
};

bool geomOk(const RefineGeom& g) {
    const int w = g.box.x1 - g.box.x0, h = g.box.y1 - g.box.y0;
    return g.factor >= 2 && g.factor <= 4 && w > 8 && h > 8 && g.box.x0 >= 2 && g.box.y0 >= 2 &&
           g.box.x1 <= g.coarse.nx - 2 && g.box.y1 <= g.coarse.ny - 2 && g.fine.nx == g.factor * w &&
           g.fine.ny == g.factor * h && g.fine.nz == g.factor * g.coarse.nz;
}

}  // namespace

int launchBandSnapshot(const LatticeView& coarse, int layout, const RefineGeom& g, const int* bandXY, int nband,
                       float* band, void* stream) {
    if (!geomOk(g) || !bandXY || !band || nband <= 0) return static_cast<int>(hipErrorInvalidValue);
    const RArgs r = makeRArgs(g);
    const hipStream_t st = static_cast<hipStream_t>(stream);
    switch (coarse.storage) {
        case kStoreF32: return RefineOps<StoreF32>::snapshot(coarse, layout, r, bandXY, nband, band, st);
        case kStoreF16: return RefineOps<StoreF16>::snapshot(coarse, layout, r, bandXY, nband, band, st);
        case kStoreBF16: return RefineOps<StoreBF16>::snapshot(coarse, layout, r, bandXY, nband, band, st);
        case kStoreF64: return RefineOps<StoreF64>::snapshot(coarse, layout, r, bandXY, nband, band, st);
        case kStoreF16C: return RefineOps<StoreF16C>::snapshot(coarse, layout, r, bandXY, nband, band, st);
        default: return static_cast<int>(hipErrorInvalidValue);
    }
}

long long fillCacheFloats(const RefineGeom& g) {
    return 39LL * (4LL * g.fine.nx + 4LL * (g.fine.ny - 4)) * g.fine.nz;
}

int launchFineFill(const LatticeView& coarse, int cLayout, const float* band, const int* bandIndex, int nband,
                   const LatticeView& fine, int fLayout, const RefineGeom& g, float timeWeight, float neqScale,
                   int fullVolume, float* cache, int cacheMode, void* stream) {
    if (!geomOk(g) || coarse.storage != fine.storage) return static_cast<int>(hipErrorInvalidValue);
    if (!fullVolume && (!band || !bandIndex || nband <= 0)) return static_cast<int>(hipErrorInvalidValue);
    const RArgs r = makeRArgs(g);
    const hipStream_t st = static_cast<hipStream_t>(stream);
    const bool full = fullVolume != 0;
    switch (coarse.storage) {
        case kStoreF32:
            return RefineOps<StoreF32>::fill(coarse, cLayout, band, bandIndex, nband, fine, fLayout, r, timeWeight,
                                             neqScale, full, cache, cacheMode, st);
        case kStoreF16:
            return RefineOps<StoreF16>::fill(coarse, cLayout, band, bandIndex, nband, fine, fLayout, r, timeWeight,
                                             neqScale, full, cache, cacheMode, st);
        case kStoreBF16:
            return RefineOps<StoreBF16>::fill(coarse, cLayout, band, bandIndex, nband, fine, fLayout, r,
                                              timeWeight, neqScale, full, cache, cacheMode, st);
        case kStoreF64:
            return RefineOps<StoreF64>::fill(coarse, cLayout, band, bandIndex, nband, fine, fLayout, r, timeWeight,
                                             neqScale, full, cache, cacheMode, st);
        case kStoreF16C:
            return RefineOps<StoreF16C>::fill(coarse, cLayout, band, bandIndex, nband, fine, fLayout, r,
                                              timeWeight, neqScale, full, cache, cacheMode, st);
        default: return static_cast<int>(hipErrorInvalidValue);
    }
}

long long restrictScratchFloats(const RefineGeom& g) {
    const long long rw = g.box.x1 - g.box.x0 - 2 * kRestrictBand, rh = g.box.y1 - g.box.y0 - 2 * kRestrictBand;
    return rw > 0 && rh > 0 ? 24LL * rw * rh * g.coarse.nz : 0;
}

int launchFineRestrict(const LatticeView& fine, int fLayout, const LatticeView& coarse, int cLayout,
                       const RefineGeom& g, float neqScaleInv, float* rho, float* u, float* v, float* w,
                       float* scratch, int recordsReady, void* stream) {
    if (!geomOk(g) || coarse.storage != fine.storage) return static_cast<int>(hipErrorInvalidValue);
    const RArgs r = makeRArgs(g);
    const hipStream_t st = static_cast<hipStream_t>(stream);
    switch (coarse.storage) {
        case kStoreF32:
            return RefineOps<StoreF32>::restrictTo(fine, fLayout, coarse, cLayout, r, neqScaleInv, rho, u, v, w, scratch, recordsReady != 0, st);
        case kStoreF16:
            return RefineOps<StoreF16>::restrictTo(fine, fLayout, coarse, cLayout, r, neqScaleInv, rho, u, v, w, scratch, recordsReady != 0, st);
        case kStoreBF16:
            return RefineOps<StoreBF16>::restrictTo(fine, fLayout, coarse, cLayout, r, neqScaleInv, rho, u, v, w,
                                                    scratch, recordsReady != 0, st);
        case kStoreF64:
            return RefineOps<StoreF64>::restrictTo(fine, fLayout, coarse, cLayout, r, neqScaleInv, rho, u, v, w, scratch, recordsReady != 0, st);
        case kStoreF16C:
            return RefineOps<StoreF16C>::restrictTo(fine, fLayout, coarse, cLayout, r, neqScaleInv, rho, u, v, w,
                                                    scratch, recordsReady != 0, st);
        default: return static_cast<int>(hipErrorInvalidValue);
    }
}

}  // namespace fc
