// Device-side helpers shared by the gfx950 kernels: wave64 DPP reductions and a
// deterministic single-pass grid reduction.
#pragma once

#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace senk_amd {

constexpr int kBlock = 256;           // threads per workgroup (4 wave64)
constexpr int kWavesPerBlock = kBlock / 64;

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

__device__ __forceinline__ double load_agent(const double *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Grid reduction of K per-thread values in one kernel: each workgroup writes its
// partial sums, the last workgroup to arrive (arrival counter, reset by that same
// workgroup) adds the partials in block order, so the result is independent of
// scheduling.  Returns true in every thread of the last workgroup, with tot[] valid
// in all of its threads.  All threads of the grid must call it.
template <int K>
__device__ __forceinline__ bool grid_reduce(const double (&v)[K], double *partials, unsigned *counter,
                                            double (&tot)[K]) {
    __shared__ double sh[K][kWavesPerBlock];
    __shared__ int is_last;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < K; k++) {
        double s = wave_sum(v[k]);
        if (lane == 0) sh[k][w] = s;
    }
    __syncthreads();
    if (threadIdx.x < K) {
        double s = 0;
#pragma unroll
        for (int i = 0; i < kWavesPerBlock; i++) s += sh[threadIdx.x][i];
        partials[(size_t)threadIdx.x * kRedMaxBlocks + blockIdx.x] = s;
    }
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned t = atomicAdd(counter, 1u);
        is_last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (!is_last) return false;
    __threadfence();
    __shared__ double shb[kWavesPerBlock];
#pragma unroll
    for (int k = 0; k < K; k++) {
        double s = 0;
        for (unsigned b = threadIdx.x; b < gridDim.x; b += blockDim.x)
            s += load_agent(partials + (size_t)k * kRedMaxBlocks + b);
        tot[k] = block_sum(s, shb);
    }
    if (threadIdx.x == 0) atomicExch(counter, 0u);
    return true;
}

// Same for a runtime count K <= KMAX (used by the multi-vector Gram-Schmidt passes).
template <int KMAX>
__device__ __forceinline__ bool grid_reduce_n(const double *v, int K, double *partials, unsigned *counter,
                                              double *tot_sh /* KMAX doubles in LDS */) {
    __shared__ double sh[KMAX][kWavesPerBlock];
    __shared__ int is_last;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < KMAX; k++) {
        if (k < K) {
            double s = wave_sum(v[k]);
            if (lane == 0) sh[k][w] = s;
        }
    }
    __syncthreads();
    if (threadIdx.x < K) {
        double s = 0;
#pragma unroll
        for (int i = 0; i < kWavesPerBlock; i++) s += sh[threadIdx.x][i];
        partials[(size_t)threadIdx.x * kRedMaxBlocks + blockIdx.x] = s;
    }
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned t = atomicAdd(counter, 1u);
        is_last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (!is_last) return false;
    __threadfence();
    __shared__ double shb[kWavesPerBlock];
    for (int k = 0; k < K; k++) {
        double s = 0;
        for (unsigned b = threadIdx.x; b < gridDim.x; b += blockDim.x)
            s += load_agent(partials + (size_t)k * kRedMaxBlocks + b);
        s = block_sum(s, shb);
        if (threadIdx.x == 0) tot_sh[k] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) atomicExch(counter, 0u);
    return true;
}

// Grid size for a grid-stride reduction pass over n elements.
inline int red_grid(long long n, int per_thread = 4) {
    long long b = (n + (long long)kBlock * per_thread - 1) / ((long long)kBlock * per_thread);
    if (b < 1) b = 1;
    if (b > kRedMaxBlocks) b = kRedMaxBlocks;
    return (int)b;
}

}  // namespace senk_amd
