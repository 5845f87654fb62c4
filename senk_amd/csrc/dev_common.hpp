// Device-side helpers shared by the gfx950 kernels: wave64 DPP reductions and
// deterministic cross-workgroup reductions through per-workgroup partials.
#pragma once

#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace senk_amd {

constexpr int kBlock = 256;           // threads per workgroup (4 wave64)
constexpr int kWavesPerBlock = kBlock / 64;
constexpr int kXcds = 8;               // MI355X: 8 accelerator dies, one L2 each

// The dispatcher hands workgroup b to XCD b % 8.  Renumbered, XCD j works through the
// j-th contiguous eighth of the grid, so that neighbouring rows -- and the vector
// entries they gather -- share one L2 instead of being fetched by every XCD.  A
// bijection on [0, gridDim.x) for any grid size.
__device__ __forceinline__ int xcd_block() {
    const int n = (int)gridDim.x, b = (int)blockIdx.x;
    const int x = b % kXcds, per = n / kXcds, rem = n % kXcds;
    return x * per + (x < rem ? x : rem) + b / kXcds;
}

// One DPP move of a double (two 32-bit halves).  Lanes whose source is outside the
// row or whose row is masked off receive 0.
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ double dpp_f64(double v) {
    int lo = __double2loint(v), hi = __double2hiint(v);
    lo = __builtin_amdgcn_update_dpp(0, lo, CTRL, ROW_MASK, 0xF, false);
    hi = __builtin_amdgcn_update_dpp(0, hi, CTRL, ROW_MASK, 0xF, false);
    return __hiloint2double(hi, lo);
}

// Sum over the 64 lanes of the wave; every lane must be active.  The total ends in
// lane 63 (quad swaps, row shifts by 4 and 8, then the row-15 and row-31 broadcasts)
// and is then read back to all lanes.
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_f64<0xB1>(v);         // quad_perm [1,0,3,2]
    v += dpp_f64<0x4E>(v);         // quad_perm [2,3,0,1]
    v += dpp_f64<0x114>(v);        // row_shr:4
    v += dpp_f64<0x118>(v);        // row_shr:8
    v += dpp_f64<0x142, 0xA>(v);   // row_bcast:15 into rows 1 and 3
    v += dpp_f64<0x143, 0xC>(v);   // row_bcast:31 into rows 2 and 3
    int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
    int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
    return __hiloint2double(hi, lo);
}

// Sum over a workgroup of kBlock threads; all threads get the total.  `sh` must hold
// kWavesPerBlock doubles and is free again on return.
__device__ __forceinline__ double block_sum(double v, double *sh) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double t = 0;
#pragma unroll
    for (int i = 0; i < kWavesPerBlock; i++) t += sh[i];
    __syncthreads();
    return t;
}

// Cross-workgroup reductions never synchronise inside a kernel: every workgroup
// stores its partial sums (store_partials), and the consumer -- the next kernel in
// stream order, or a small finalize kernel -- adds the partials in a fixed order
// (sum_partials).  Kernel boundaries make the partials visible, so no device-scope
// fence is needed (on MI355X such a fence writes back the per-XCD L2), and the result
// does not depend on scheduling.
constexpr int kPartLd = kRedMaxBlocks;  // stride between the K partial arrays of a slot

template <int K>
__device__ __forceinline__ void store_partials(const double (&v)[K], double *partials) {
    __shared__ double sh[K][kWavesPerBlock];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const double s = wave_sum(v[k]);
        if (lane == 0) sh[k][w] = s;
    }
    __syncthreads();
    if (threadIdx.x < K) {
        double s = 0;
#pragma unroll
        for (int i = 0; i < kWavesPerBlock; i++) s += sh[threadIdx.x][i];
        partials[(size_t)threadIdx.x * kPartLd + blockIdx.x] = s;
    }
}
// Runtime count K <= KMAX.
template <int KMAX>
__device__ __forceinline__ void store_partials_n(const double *v, int K, double *partials) {
    __shared__ double sh[KMAX][kWavesPerBlock];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < KMAX; k++) {
        if (k < K) {
            const double s = wave_sum(v[k]);
            if (lane == 0) sh[k][w] = s;
        }
    }
    __syncthreads();
    if (threadIdx.x < K) {
        double s = 0;
#pragma unroll
        for (int i = 0; i < kWavesPerBlock; i++) s += sh[threadIdx.x][i];
        partials[(size_t)threadIdx.x * kPartLd + blockIdx.x] = s;
    }
}
// Sum of p[0..B) in a fixed order; every thread of the workgroup gets the total.  All
// threads must call it.
__device__ __forceinline__ double sum_partials(const double *p, int B) {
    __shared__ double sh[kWavesPerBlock];
    double s = 0;
    for (int b = threadIdx.x; b < B; b += kBlock) s += p[b];
    return block_sum(s, sh);
}

// ----------------------------------------------------------------- 16-byte streaming
// Streaming passes move 16 bytes per lane per access (a float4, or two double2): on
// MI355X that streams at ~6.4 TB/s where 4- and 8-byte accesses reach 2.5-4.3 TB/s.
// Each thread takes groups of four consecutive elements (SENK_FOR4); a group that runs
// past n is loaded element by element with zero fill and stored masked, and so is every
// group of a vector whose base is not 16-byte aligned (work vectors are aligned, and
// the columns of a basis are stored at a stride that is a multiple of 4 elements, pad4).
__host__ __device__ inline size_t pad4(size_t n) { return (n + 3) & ~(size_t)3; }
__device__ __forceinline__ void ld4(const float *p, float (&o)[4]) {
    const float4 t = *reinterpret_cast<const float4 *>(p);
    o[0] = t.x; o[1] = t.y; o[2] = t.z; o[3] = t.w;
}
__device__ __forceinline__ void ld4(const double *p, double (&o)[4]) {
    const double2 a = *reinterpret_cast<const double2 *>(p), b = *reinterpret_cast<const double2 *>(p + 2);
    o[0] = a.x; o[1] = a.y; o[2] = b.x; o[3] = b.y;
}
__device__ __forceinline__ void st4(float *p, const float (&o)[4]) {
    *reinterpret_cast<float4 *>(p) = make_float4(o[0], o[1], o[2], o[3]);
}
__device__ __forceinline__ void st4(double *p, const double (&o)[4]) {
    *reinterpret_cast<double2 *>(p) = make_double2(o[0], o[1]);
    *reinterpret_cast<double2 *>(p + 2) = make_double2(o[2], o[3]);
}
__device__ __forceinline__ bool aligned16(const void *p) { return ((uintptr_t)p & 15) == 0; }
template <typename T>
__device__ __forceinline__ void get4(const T *p, long long i0, long long n, T (&o)[4]) {
    if (i0 + 4 <= n && aligned16(p)) {
        ld4(p + i0, o);
    } else {
#pragma unroll
        for (int u = 0; u < 4; u++) o[u] = i0 + u < n ? p[i0 + u] : T(0);
    }
}
// get4 with `fill` in the places at or beyond n (1 for a divisor: a zero there would turn
// the group's zero terms into 0 * inf)
template <typename T>
__device__ __forceinline__ void get4_fill(const T *p, long long i0, long long n, T (&o)[4], T fill) {
    if (i0 + 4 <= n && aligned16(p)) {
        ld4(p + i0, o);
    } else {
#pragma unroll
        for (int u = 0; u < 4; u++) o[u] = i0 + u < n ? p[i0 + u] : fill;
    }
}
template <typename T>
__device__ __forceinline__ void put4(T *p, long long i0, long long n, const T (&o)[4]) {
    if (i0 + 4 <= n && aligned16(p)) {
        st4(p + i0, o);
    } else {
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (i0 + u < n) p[i0 + u] = o[u];
    }
}
#define SENK_FOR4(i0, n)                                                                        \
    for (long long i0 = ((long long)blockIdx.x * kBlock + threadIdx.x) * 4; i0 < (long long)(n); \
         i0 += (long long)gridDim.x * kBlock * 4)

// Streaming passes: kUnroll elements per thread per trip, at most kRedMaxBlocks
// workgroups (so a consumer sums at most that many partials).
constexpr int kUnroll = 4;
inline int red_grid(long long n) {
    long long b = (n + (long long)kBlock * kUnroll - 1) / ((long long)kBlock * kUnroll);
    if (b < 1) b = 1;
    if (b > kRedMaxBlocks) b = kRedMaxBlocks;
    return (int)b;
}
// for (i over this thread's elements) with kUnroll independent loads in flight
#define SENK_STREAM(i, n, ...)                                                                          \
    for (long long i##0_ = (long long)blockIdx.x * kBlock * kUnroll + threadIdx.x; i##0_ < (n);       \
         i##0_ += (long long)gridDim.x * kBlock * kUnroll) {                                            \
        _Pragma("unroll") for (int u_ = 0; u_ < kUnroll; u_++) {                                        \
            const long long i = i##0_ + (long long)u_ * kBlock;                                         \
            if (i < (n)) { __VA_ARGS__ }                                                                \
        }                                                                                               \
    }

}  // namespace senk_amd
