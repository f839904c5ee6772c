// Vortex-generator geometry (see fc_vg.hpp). Compiled with -ffp-contract=off
// so the stamped flag fields follow the same FP32 arithmetic as the reference.
#include "fc_vg.hpp"

#include <algorithm>
#include <cmath>
#include <limits>

namespace fc {
namespace {

constexpr float kPi = 3.14159265358979f;

inline Vec2 sub(Vec2 a, Vec2 b) { return Vec2{a.x - b.x, a.y - b.y}; }
inline Vec2 add(Vec2 a, Vec2 b) { return Vec2{a.x + b.x, a.y + b.y}; }
inline Vec2 mul(Vec2 a, float s) { return Vec2{a.x * s, a.y * s}; }
inline Vec2 div(Vec2 a, float s) { return Vec2{a.x / s, a.y / s}; }
inline float dot(Vec2 a, Vec2 b) { return a.x * b.x + a.y * b.y; }
inline Vec2 normalized(Vec2 v) {
    const float l = std::sqrt(dot(v, v));
    return l > 1e-12f ? div(v, l) : v;
}
inline Vec2 perpCCW(Vec2 v) { return Vec2{-v.y, v.x}; }
inline Vec2 rotated(Vec2 v, float r) {
    const float c = std::cos(r);
    const float s = std::sin(r);
    return Vec2{v.x * c - v.y * s, v.x * s + v.y * c};
}

inline Vec3 add(Vec3 a, Vec3 b) { return Vec3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline Vec3 sub(Vec3 a, Vec3 b) { return Vec3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline Vec3 mul(Vec3 a, float s) { return Vec3{a.x * s, a.y * s, a.z * s}; }
inline float dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Vec3 cross(Vec3 a, Vec3 b) {
    return Vec3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline Vec3 normalized(Vec3 v) {
    const float l = std::sqrt(dot(v, v));
    return l > 1e-12f ? Vec3{v.x / l, v.y / l, v.z / l} : v;
}

inline long long cellIndex(int x, int y, int z, const Dims& d) {
    return static_cast<long long>(x) +
           static_cast<long long>(d.nx) * (static_cast<long long>(y) + static_cast<long long>(d.ny) * z);
}

// Fluid -> Solid at (x, y, z) with z periodic; face planes are never touched.
inline void stampCell(std::vector<std::uint8_t>& flags, const Dims& d, int x, int y, int z) {
    if (x < 1 || x > d.nx - 2 || y < 1 || y > d.ny - 2) return;
    const int zw = ((z % d.nz) + d.nz) % d.nz;
    const std::size_t i = static_cast<std::size_t>(cellIndex(x, y, zw, d));
    if (flags[i] == kFluid) flags[i] = kSolid;
}

// One vane or ramp wedge: slices along the vane axis, each slice's root on
// the local suction surface (embedded 1.5 cells), rasterised by 0.45-cell
// sampling in height and thickness (vg.cpp:78-171).
void stampVane(const Airfoil& foil, float aoaRad, const DomainLayout& layout, std::vector<std::uint8_t>& flags,
               float x_c, float zCenter, float betaRad, float hCells, float lenCells, float halfW, bool ramp,
               std::vector<VaneSlab>* slabs) {
    const SurfaceFrame center = surfaceFrameAt(foil, x_c, true);
    if (!center.valid) return;
    const Vec2 qc{0.25f, 0.0f};
    if (slabs) {
        const float chordC = static_cast<float>(layout.chordCells);
        const Vec2 cpr = rotated(sub(center.point, qc), -aoaRad);
        const Vec2 cnr = rotated(center.normal, -aoaRad);
        const Vec2 ctr = rotated(center.tangent, -aoaRad);
        const Vec3 root{layout.anchorX() + cpr.x * chordC, layout.anchorY() + cpr.y * chordC, zCenter};
        const Vec3 n3{cnr.x, cnr.y, 0.0f};
        const Vec3 d3{std::cos(betaRad) * ctr.x, std::cos(betaRad) * ctr.y, std::sin(betaRad)};
        const Vec3 b3 = normalized(cross(n3, d3));
        const float embed = 1.5f;
        VaneSlab s;
        s.center = add(root, mul(n3, 0.5f * (hCells - embed)));
        s.nAxis = normalized(n3);
        s.dAxis = normalized(d3);
        s.bAxis = b3;
        s.halfHeight = 0.5f * (hCells + embed);
        s.halfLength = 0.5f * lenCells;
        s.halfThick = halfW;
        slabs->push_back(s);
    }
    const float cosB = std::cos(betaRad);
    const float sinB = std::sin(betaRad);
    const float chord = static_cast<float>(layout.chordCells);
    const float ax = layout.anchorX();
    const float ay = layout.anchorY();
    const float kStep = 0.45f;
    const float kEmbed = 1.5f;
    for (float s = -0.5f * lenCells; s <= 0.5f * lenCells + 1e-4f; s += kStep) {
        const float station = std::clamp(x_c + (s * cosB * center.tangent.x) / chord, 0.0f, 1.0f);
        const SurfaceFrame f = surfaceFrameAt(foil, station, true);
        if (!f.valid) continue;
        const Vec2 pr = rotated(sub(f.point, qc), -aoaRad);
        const Vec2 nr = rotated(f.normal, -aoaRad);
        const Vec2 tr = rotated(f.tangent, -aoaRad);
        const Vec3 root{ax + pr.x * chord, ay + pr.y * chord, zCenter + s * sinB};
        const Vec3 n3{nr.x, nr.y, 0.0f};
        const Vec3 d3{cosB * tr.x, cosB * tr.y, sinB};
        const Vec3 b3 = normalized(cross(n3, d3));
        const float top = ramp ? hCells * (s + 0.5f * lenCells) / std::max(lenCells, 1e-3f) : hCells;
        for (float u = -kEmbed; u <= top + 1e-4f; u += kStep) {
            for (float w = -halfW; w <= halfW + 1e-4f; w += kStep) {
                const Vec3 p = add(add(root, mul(n3, u)), mul(b3, w));
                stampCell(flags, layout.dims, static_cast<int>(std::floor(p.x)), static_cast<int>(std::floor(p.y)),
                          static_cast<int>(std::floor(p.z)));
            }
        }
    }
}

}  // namespace

SurfaceFrame surfaceFrameAt(const Airfoil& foil, float x_c, bool upper) {
    SurfaceFrame fr;
    if (!foil.valid() || x_c < 0.0f || x_c > 1.0f) return fr;
    const int n = static_cast<int>(foil.points.size());
    if (foil.leIndex <= 0 || foil.leIndex >= n - 1) return fr;
    const int begin = upper ? 0 : foil.leIndex;
    const int end = upper ? foil.leIndex : n - 1;
    int best = -1;
    float bestD = std::numeric_limits<float>::max();
    for (int i = begin; i < end; ++i) {
        const Vec2& a = foil.points[static_cast<std::size_t>(i)];
        const Vec2& b = foil.points[static_cast<std::size_t>(i + 1)];
        const float lo = std::min(a.x, b.x);
        const float hi = std::max(a.x, b.x);
        const float dd = (x_c < lo) ? (lo - x_c) : (x_c > hi ? x_c - hi : 0.0f);
        if (dd < bestD) {
            bestD = dd;
            best = i;
            if (dd == 0.0f && hi - lo > 1e-9f) break;
        }
    }
    if (best < 0) return fr;
    const Vec2& a = foil.points[static_cast<std::size_t>(best)];
    const Vec2& b = foil.points[static_cast<std::size_t>(best + 1)];
    const float dx = b.x - a.x;
    const float t = (std::fabs(dx) > 1e-9f) ? std::clamp((x_c - a.x) / dx, 0.0f, 1.0f) : 0.5f;
    fr.point = add(a, mul(sub(b, a), t));
    fr.tangent = normalized(upper ? sub(a, b) : sub(b, a));
    fr.normal = upper ? perpCCW(fr.tangent) : mul(perpCCW(fr.tangent), -1.0f);
    fr.valid = true;
    return fr;
}

void voxelizeVG(const VGParams& vg, const Airfoil& foil, float aoaDeg, const DomainLayout& layout,
                std::vector<std::uint8_t>& flags, std::vector<VaneSlab>* slabs) {
    if (!vg.enabled) return;
    if (!foil.valid() || layout.dims.nz < 1) return;
    const float h = std::max(1.0f, vg.height_c * static_cast<float>(layout.chordCells));
    const float len = std::max(1.0f, vg.length_h * h);
    const float thick = std::clamp(h / 6.0f, 1.0f, 2.0f);
    const float beta = vg.beta_deg * kPi / 180.0f;
    const float aoaRad = aoaDeg * kPi / 180.0f;
    const float pitch = vg.pitch_c * static_cast<float>(layout.chordCells);
    const int count = std::max(1, vg.count);
    const float x_c = std::clamp(vg.x_c, 0.0f, 1.0f);
    const float zMid = 0.5f * static_cast<float>(layout.dims.nz);
    for (int i = 0; i < count; ++i) {
        const float zc = zMid + (static_cast<float>(i) - 0.5f * static_cast<float>(count - 1)) * pitch;
        switch (vg.type) {
            case 0:
            case 2:
                stampVane(foil, aoaRad, layout, flags, x_c, zc, beta, h, len, 0.5f * thick, false, slabs);
                break;
            case 1: {
                const float halfGap = 0.5f * vg.gap_h * h;
                const float bn = vg.commonFlowDown ? beta : -beta;
                stampVane(foil, aoaRad, layout, flags, x_c, zc - halfGap, bn, h, len, 0.5f * thick, false, slabs);
                stampVane(foil, aoaRad, layout, flags, x_c, zc + halfGap, -bn, h, len, 0.5f * thick, false, slabs);
                break;
            }
            case 3:
                stampVane(foil, aoaRad, layout, flags, x_c, zc, beta, h, len, std::max(0.5f, 0.5f * h), true,
                          nullptr);
                break;
            default:
                break;
        }
    }
}

std::vector<std::uint8_t> buildFlagsWithVGs(const std::vector<VGParams>& vgs, const Airfoil& foil, float aoaDeg,
                                            const DomainLayout& layout, const std::vector<std::uint8_t>& clean,
                                            std::vector<VaneSlab>* slabs) {
    std::vector<std::uint8_t> flags = clean;
    for (const VGParams& vg : vgs) voxelizeVG(vg, foil, aoaDeg, layout, flags, slabs);
    closeTrailingEdgeGaps(layout.dims, flags);
    return flags;
}

float rayOBBEntry(const Vec3& origin, const Vec3& dir, const VaneSlab& slab) {
    const Vec3 p = sub(origin, slab.center);
    const Vec3 axes[3] = {slab.nAxis, slab.dAxis, slab.bAxis};
    const float halfs[3] = {slab.halfHeight, slab.halfLength, slab.halfThick};
    float tMin = 0.0f, tMax = 1.0f;
    for (int a = 0; a < 3; ++a) {
        const float e = dot(p, axes[a]);
        const float f = dot(dir, axes[a]);
        const float h = halfs[a];
        if (std::fabs(f) < 1e-7f) {
            if (-e - h > 0.0f || -e + h < 0.0f) return -1.0f;
        } else {
            float t1 = (-e - h) / f;
            float t2 = (-e + h) / f;
            if (t1 > t2) std::swap(t1, t2);
            tMin = std::max(tMin, t1);
            tMax = std::min(tMax, t2);
            if (tMin > tMax) return -1.0f;
        }
    }
    return tMin;
}

QLinkList buildVaneQLinks(const Dims& d, const std::vector<std::uint8_t>& active,
                          const std::vector<std::uint8_t>& clean, const std::vector<VaneSlab>& slabs) {
    QLinkList out;
    if (slabs.empty() || active.size() != clean.size()) return out;
    auto solidAt = [&](const std::vector<std::uint8_t>& f, int x, int y, int z) {
        if (x < 0 || x >= d.nx || y < 0 || y >= d.ny) return false;
        const int zw = ((z % d.nz) + d.nz) % d.nz;
        return f[static_cast<std::size_t>(cellIndex(x, y, zw, d))] == kSolid;
    };
    auto flagAt = [&](int x, int y, int z) -> std::uint8_t {
        if (x < 0 || x >= d.nx || y < 0 || y >= d.ny) return kSolid;
        const int zw = ((z % d.nz) + d.nz) % d.nz;
        return active[static_cast<std::size_t>(cellIndex(x, y, zw, d))];
    };
    for (int z = 0; z < d.nz; ++z)
        for (int y = 0; y < d.ny; ++y)
            for (int x = 0; x < d.nx; ++x) {
                const long long cell = cellIndex(x, y, z, d);
                if (active[static_cast<std::size_t>(cell)] != kFluid) continue;
                for (int q = 1; q < kQ; ++q) {
                    const int sx = x - kCx[q], sy = y - kCy[q], sz = z - kCz[q];
                    if (!solidAt(active, sx, sy, sz) || solidAt(clean, sx, sy, sz)) continue;  // vane solid only
                    const Vec3 origin{static_cast<float>(x) + 0.5f, static_cast<float>(y) + 0.5f,
                                      static_cast<float>(z) + 0.5f};
                    const Vec3 dir{static_cast<float>(-kCx[q]), static_cast<float>(-kCy[q]),
                                   static_cast<float>(-kCz[q])};
                    float qf = -1.0f;
                    for (const VaneSlab& s : slabs) {
                        const float t = rayOBBEntry(origin, dir, s);
                        if (t >= 0.0f && (qf < 0.0f || t < qf)) qf = t;
                    }
                    if (qf < 0.0f) continue;
                    if (qf < kQLibbMin || qf > 1.0f) continue;
                    const std::uint8_t ffFlag = flagAt(x + kCx[q], y + kCy[q], z + kCz[q]);
                    if (ffFlag == kInterface) continue;
                    out.cellIdx.push_back(cell);
                    out.dir.push_back(static_cast<std::uint8_t>(q));
                    out.ffFluid.push_back(ffFlag == kFluid ? 1u : 0u);
                    out.q.push_back(qf);
                }
            }
    return out;
}

void densifyQLinks(const QLinkList& links, long long ncells, std::vector<std::uint8_t>& qFrac,
                   std::vector<std::uint32_t>& ffMask) {
    qFrac.assign(static_cast<std::size_t>(ncells) * kQ, 0);
    ffMask.assign(static_cast<std::size_t>(ncells), 0u);
    for (std::size_t i = 0; i < links.size(); ++i) {
        const long long cell = links.cellIdx[i];
        const int q = links.dir[i];
        const float qc = std::clamp(links.q[i], kQLibbMin, 1.0f);
        std::uint8_t b = static_cast<std::uint8_t>(std::lround(qc * 255.0f));
        if (b == 0) b = 1;
        qFrac[static_cast<std::size_t>(cell) * kQ + q] = b;
        if (links.ffFluid[i]) ffMask[static_cast<std::size_t>(cell)] |= (1u << q);
    }
}

}  // namespace fc
