// Interpolated bounce-back (q-LIBB) on the in-place population store.
//
// The reference applies the sub-cell wall position inside its pull loop
// (src/sim/lbm_core.cu:338-364, formula as written there): for a fluid cell c
// whose pull of direction q hits a solid with cut byte b != 0 and qf = b/255,
//   qf >= 1/2          : f_q = (1/(2qf)) f*_q(c) + (1 - 1/(2qf)) f*_opp(q)(c)
//   qf <  1/2, ff bit  : f_q = 2qf f*_q(c) + (1 - 2qf) f*_q(c + e_q)
//   otherwise          : plain half-way value f*_opp(q)(c).
// In place, f*_q(c) and f*_q(c + e_q) are overwritten during the step, so
// this pass runs between steps over the sparse link list: the gather kernel
// reads the three previous-step populations of every link, the scatter kernel
// then writes the interpolated value into the link's bounce-back slot (read by
// exactly that link in the next step). All three populations carry the same
// weight, so the blend is done on the stored deviations directly.
#include <hip/hip_runtime.h>

#include "fc_kcommon.hpp"

namespace fc {
namespace {

// Slot of the post-collision population f*_q of cell c in layout `layout`.
__device__ inline long long slotOf(const KArgs& a, int layout, int q, int x, int y, int z) {
    if (layout == kLayoutL) {
        return static_cast<long long>(kOpp[q]) * a.ncells + x +
               static_cast<long long>(a.nx) * (y + static_cast<long long>(a.ny) * z);
    }
    int xx = x + kCx[q], yy = y + kCy[q], zz = z + kCz[q];
    if (zz < 0) zz += a.nz;
    if (zz >= a.nz) zz -= a.nz;
    if (xx < 0) xx += a.nx;
    if (xx >= a.nx) xx -= a.nx;
    if (yy < 0) yy += a.ny;
    if (yy >= a.ny) yy -= a.ny;
    return static_cast<long long>(q) * a.ncells + xx + static_cast<long long>(a.nx) * (yy + static_cast<long long>(a.ny) * zz);
}

template <class S>
__global__ void __launch_bounds__(256) gatherKernel(KArgs a, int layout, const int* cells,
                                                    const std::uint32_t* packed, int count, double* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    using T = typename S::T;
    using C = typename S::C;
    const T* f = static_cast<const T*>(a.f);
    const int c = cells[i];
    const std::uint32_t pk = packed[i];
    const int q = static_cast<int>(pk & 0xFFu);
    const int qb = static_cast<int>((pk >> 8) & 0xFFu);
    const bool ff = (pk >> 16) & 1u;
    const int nxny = a.nx * a.ny;
    const int z = c / nxny, rem = c - z * nxny, y = rem / a.nx, x = rem - y * a.nx;
    // half-way source f*_opp(q)(c) and the outgoing f*_q(c)
    const C hBB = S::load(f + slotOf(a, layout, kOpp[q], x, y, z));
    const C hOut = S::load(f + slotOf(a, layout, q, x, y, z));
    const C qf = C(static_cast<float>(qb) * (1.0f / 255.0f));
    C v = hBB;
    if (qf >= C(0.5)) {
        const C inv2q = C(0.5) / qf;
        v = inv2q * hOut + (C(1) - inv2q) * hBB;
    } else if (ff) {
        int xf = x + kCx[q], yf = y + kCy[q], zf = z + kCz[q];
        if (zf < 0) zf += a.nz;
        if (zf >= a.nz) zf -= a.nz;
        if (xf < 0) xf += a.nx;
        if (xf >= a.nx) xf -= a.nx;
        if (yf < 0) yf += a.ny;
        if (yf >= a.ny) yf -= a.ny;
        const C hFF = S::load(f + slotOf(a, layout, q, xf, yf, zf));
        v = C(2) * qf * hOut + (C(1) - C(2) * qf) * hFF;
    }
    out[i] = static_cast<double>(v);
}

template <class S>
__global__ void __launch_bounds__(256) scatterKernel(KArgs a, int layout, const int* cells,
                                                     const std::uint32_t* packed, int count, const double* vals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    using T = typename S::T;
    using C = typename S::C;
    T* f = static_cast<T*>(a.f);
    const int c = cells[i];
    const int q = static_cast<int>(packed[i] & 0xFFu);
    const int nxny = a.nx * a.ny;
    const int z = c / nxny, rem = c - z * nxny, y = rem / a.nx, x = rem - y * a.nx;
    S::store(f + slotOf(a, layout, kOpp[q], x, y, z), static_cast<C>(vals[i]));
}

template <class S>
int runQ(const LatticeView& lat, int layout, const int* cells, const std::uint32_t* packed, int count,
         double* scratch, hipStream_t s) {
    KArgs a = makeArgs(lat);
    const unsigned blocks = static_cast<unsigned>((count + 255) / 256);
    gatherKernel<S><<<blocks, 256, 0, s>>>(a, layout, cells, packed, count, scratch);
    scatterKernel<S><<<blocks, 256, 0, s>>>(a, layout, cells, packed, count, scratch);
    return static_cast<int>(hipGetLastError());
}

}  // namespace

int launchQLibbPrepass(const LatticeView& lat, int layout, const int* cells, const std::uint32_t* packed,
                       int count, double* scratch, void* stream) {
    if (count <= 0) return 0;
    if (!cells || !packed || !scratch) return static_cast<int>(hipErrorInvalidValue);
    const hipStream_t s = static_cast<hipStream_t>(stream);
    switch (lat.storage) {
        case kStoreF32: return runQ<StoreF32>(lat, layout, cells, packed, count, scratch, s);
        case kStoreF16: return runQ<StoreF16>(lat, layout, cells, packed, count, scratch, s);
        case kStoreBF16: return runQ<StoreBF16>(lat, layout, cells, packed, count, scratch, s);
        case kStoreF16C: return runQ<StoreF16C>(lat, layout, cells, packed, count, scratch, s);
        case kStoreF64: return runQ<StoreF64>(lat, layout, cells, packed, count, scratch, s);
        default: return static_cast<int>(hipErrorInvalidValue);
    }
}

}  // namespace fc
