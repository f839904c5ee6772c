// Traffic-only probes of the HBM ceiling that bounds the stream-collide
// kernels (measurement infrastructure; no counterpart in the reference).
//
// aaTrafficKernel moves exactly the bytes of one in-place (AA) step with the
// step kernels' own decomposition (one wave per 64-cell row segment, the
// XCD slab remap, 19 direction slabs): every lane reads its 19 slots and
// writes 19 slots, either all in its own cell (the S->L step: read (q, c),
// write (opp q, c)), shifted to the 18 neighbours (the L->S step: read
// (opp q, c - e_q), write (q, c + e_q)), or in the "esoteric twist" pattern
// (Geier and Schoenherr, Computation 5 (2017) 19), where every step reads and
// writes slot q at c + max(e_q, 0) componentwise: the cell itself and its
// six positive neighbours x+1, y+1, z+1 and their pairs, each slot read and
// written by exactly one lane. All directions wrap periodically, so every
// address lies inside the array. An optional `work` count adds that
// many FP32 fused multiply-adds per population (19 independent chains)
// between the read and the write, to see where arithmetic starts to cost
// bandwidth. The
// copy kernel is the plain streaming reference (one 16-byte vector per lane).
#include <hip/hip_runtime.h>

#include "fc_kcommon.hpp"
#include "../../include/foilcfd.h"

namespace fc {
namespace {

template <class T>
__device__ inline float asWork(T v) {
    return static_cast<float>(v);
}
template <class T>
__device__ inline T fromWork(float x) {
    return static_cast<T>(static_cast<int>(x));
}
template <>
__device__ inline float asWork<float>(float v) {
    return v;
}
template <>
__device__ inline float fromWork<float>(float x) {
    return x;
}

enum Pattern { kLocal = 0, kShifted = 1, kTwist = 2 };

constexpr int posPart(int c) { return c > 0 ? c : 0; }

template <class T, int PAT>
__global__ void __launch_bounds__(kBlock) aaTrafficKernel(KArgs a, int work) {
    RowPos p;
    if (!rowPos<true, true>(a, p)) return;
    T* const f = static_cast<T*>(a.f);
    T v[kQ];
    staticFor<kQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        long long i;
        if constexpr (PAT == kShifted)
            i = static_cast<long long>(kOpp[q]) * a.ncells + nb(p, -kCx[q], -kCy[q], -kCz[q]);
        else if constexpr (PAT == kTwist)
            i = static_cast<long long>(q) * a.ncells + nb(p, posPart(kCx[q]), posPart(kCy[q]), posPart(kCz[q]));
        else
            i = static_cast<long long>(q) * a.ncells + nb(p, 0, 0, 0);
        v[q] = f[i];
    });
    if (work > 0) {
        float x[kQ], s = 0.0f;
#pragma unroll
        for (int q = 0; q < kQ; ++q) {
            x[q] = asWork<T>(v[q]);
            s += x[q];
        }
        s *= 1e-6f;
        // `work` rounds of 19 independent FMAs (the collision's shape: many
        // independent chains, not one long dependent one)
        for (int k = 0; k < work; ++k)
#pragma unroll
            for (int q = 0; q < kQ; ++q) x[q] = __builtin_fmaf(x[q], 0.999f, s);
#pragma unroll
        for (int q = 0; q < kQ; ++q) v[q] = fromWork<T>(x[q]);
    }
    staticFor<kQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        long long i;
        if constexpr (PAT == kShifted)
            i = static_cast<long long>(q) * a.ncells + nb(p, kCx[q], kCy[q], kCz[q]);
        else if constexpr (PAT == kTwist)
            i = static_cast<long long>(q) * a.ncells + nb(p, posPart(kCx[q]), posPart(kCy[q]), posPart(kCz[q]));
        else
            i = static_cast<long long>(kOpp[q]) * a.ncells + nb(p, 0, 0, 0);
        f[i] = v[q];
    });
}

__global__ void __launch_bounds__(kBlock) copyKernel(const float4* __restrict__ src, float4* __restrict__ dst,
                                                     long long n) {
    const long long i = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// The step kernels' 16-bit store codec applied to given values and random
// bits (the value is the codec's input: the scaled deviation K (f - w)).
template <class S>
__global__ void __launch_bounds__(kBlock) roundKernel(const float* __restrict__ v, const unsigned* __restrict__ r,
                                                      int sto, long long n, unsigned short* __restrict__ out) {
    const long long i = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n) out[i] = roundedBits<S>(v[i], sto != 0, r[i]);
}

template <class S>
int runRound(const float* v, const unsigned* r, int sto, long long n, unsigned short* out) {
    float* dv = nullptr;
    unsigned* dr = nullptr;
    unsigned short* dout = nullptr;
    const std::size_t m = static_cast<std::size_t>(n);
    int rc = FC_OK;
    if (hipMalloc(&dv, m * 4) != hipSuccess || hipMalloc(&dr, m * 4) != hipSuccess ||
        hipMalloc(&dout, m * 2) != hipSuccess)
        rc = FC_ERR_ALLOC;
    if (rc == FC_OK && (hipMemcpy(dv, v, m * 4, hipMemcpyHostToDevice) != hipSuccess ||
                        hipMemcpy(dr, r, m * 4, hipMemcpyHostToDevice) != hipSuccess))
        rc = FC_ERR_DEVICE;
    if (rc == FC_OK) {
        roundKernel<S><<<static_cast<unsigned>((n + kBlock - 1) / kBlock), kBlock>>>(dv, dr, sto, n, dout);
        if (hipGetLastError() != hipSuccess || hipMemcpy(out, dout, m * 2, hipMemcpyDeviceToHost) != hipSuccess)
            rc = FC_ERR_DEVICE;
    }
    (void)hipFree(dv);
    (void)hipFree(dr);
    (void)hipFree(dout);
    return rc;
}

// Milliseconds per launch of `launch`, averaged over reps after 3 warm-up
// launches, bracketed by HIP events on the null stream.
template <class L>
int timeLaunches(L launch, int reps, double* ms) {
    for (int i = 0; i < 3; ++i) launch();
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return FC_ERR_DEVICE;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return FC_ERR_DEVICE;
    }
    (void)hipEventRecord(e0, nullptr);
    for (int i = 0; i < reps; ++i) launch();
    (void)hipEventRecord(e1, nullptr);
    const hipError_t se = hipEventSynchronize(e1);
    float t = 0.0f;
    const hipError_t te = hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (se != hipSuccess || te != hipSuccess || hipGetLastError() != hipSuccess) return FC_ERR_DEVICE;
    *ms = static_cast<double>(t) / reps;
    return FC_OK;
}

template <class T>
int runAA(fc_dims d, int pattern, int work, int reps, double* ms) {
    LatticeView lat;
    lat.dims = Dims{d.nx, d.ny, d.nz};
    const std::size_t bytes = static_cast<std::size_t>(kQ) * static_cast<std::size_t>(lat.dims.cells()) * sizeof(T);
    void* f = nullptr;
    if (hipMalloc(&f, bytes) != hipSuccess) return FC_ERR_ALLOC;
    int rc = hipMemset(f, 0, bytes) == hipSuccess ? FC_OK : FC_ERR_DEVICE;
    lat.f = f;
    KArgs a = makeArgs(lat);
    const dim3 g(static_cast<unsigned>(a.chunk) * kXcds), b(kBlock);
    if (rc == FC_OK) {
        if (pattern == kShifted)
            rc = timeLaunches([&] { aaTrafficKernel<T, kShifted><<<g, b>>>(a, work); }, reps, ms);
        else if (pattern == kTwist)
            rc = timeLaunches([&] { aaTrafficKernel<T, kTwist><<<g, b>>>(a, work); }, reps, ms);
        else
            rc = timeLaunches([&] { aaTrafficKernel<T, kLocal><<<g, b>>>(a, work); }, reps, ms);
    }
    (void)hipFree(f);
    return rc;
}

}  // namespace
}  // namespace fc

extern "C" {

int fc_probe_aa_traffic(fc_dims dims, int elem_bytes, int pattern, int work, int reps, double* ms_per_launch) {
    if (!ms_per_launch || dims.nx < 3 || dims.ny < 3 || dims.nz < 3 || reps < 1 || work < 0 || pattern < 0 ||
        pattern > fc::kTwist ||
        static_cast<long long>(dims.nx) * dims.ny * dims.nz >= (1ll << 31))
        return FC_ERR_INVALID;
    if (elem_bytes == 2) return fc::runAA<unsigned short>(dims, pattern, work, reps, ms_per_launch);
    if (elem_bytes == 4) return fc::runAA<float>(dims, pattern, work, reps, ms_per_launch);
    return FC_ERR_INVALID;
}

int fc_probe_store_rounding(int storage, const float* values, const unsigned* random_bits, int stochastic,
                            long long n, unsigned short* codes) {
    if (!values || !random_bits || !codes || n < 1 || n > (1ll << 30)) return FC_ERR_INVALID;
    if (storage == FC_STORAGE_F16C) return fc::runRound<fc::StoreF16C>(values, random_bits, stochastic, n, codes);
    if (storage == FC_STORAGE_F16) return fc::runRound<fc::StoreF16>(values, random_bits, stochastic, n, codes);
    if (storage == FC_STORAGE_BF16) return fc::runRound<fc::StoreBF16>(values, random_bits, stochastic, n, codes);
    return FC_ERR_INVALID;
}

int fc_probe_copy(size_t bytes, int reps, double* ms_per_launch) {
    if (!ms_per_launch || bytes < 16 || reps < 1) return FC_ERR_INVALID;
    const long long n = static_cast<long long>(bytes / 16);
    float4 *s = nullptr, *d = nullptr;
    if (hipMalloc(&s, static_cast<std::size_t>(n) * 16) != hipSuccess) return FC_ERR_ALLOC;
    if (hipMalloc(&d, static_cast<std::size_t>(n) * 16) != hipSuccess) {
        (void)hipFree(s);
        return FC_ERR_ALLOC;
    }
    int rc = (hipMemset(s, 0, static_cast<std::size_t>(n) * 16) == hipSuccess &&
              hipMemset(d, 0, static_cast<std::size_t>(n) * 16) == hipSuccess)
                 ? FC_OK
                 : FC_ERR_DEVICE;
    const dim3 g(static_cast<unsigned>((n + fc::kBlock - 1) / fc::kBlock)), b(fc::kBlock);
    if (rc == FC_OK) rc = fc::timeLaunches([&] { fc::copyKernel<<<g, b>>>(s, d, n); }, reps, ms_per_launch);
    (void)hipFree(s);
    (void)hipFree(d);
    return rc;
}

}  // extern "C"
