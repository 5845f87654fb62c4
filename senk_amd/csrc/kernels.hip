// gfx950 kernels for the SenK hot path: SpMV (CSR, SELL-C(-sigma), BCSR), colored
// SpTRSV (CSR and BCSR), ILU(0) numeric factorization per color, and BLAS-1.
//
// Floating point is compiled without contraction (-ffp-contract=off), so a kernel
// that keeps the reference's per-row operation order (SpTRSV, ILU(0), SELL SpMV) gives
// the same bits as the reference's sequential loop.
#include "dev_common.hpp"

namespace senk_amd {

// =============================================================== SpMV CSR
// G lanes per row (G | 64).  Lane q of a row sums entries rptr[row]+q, +G, ... in
// order; the G partial sums are then combined by a butterfly inside the lane group.
// Rows are walked grid-stride so that the fused residual form can reduce ||y||^2
// with a bounded grid.
template <typename T, int G>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, G);
    return v;
}

template <typename T, int G>
__global__ __launch_bounds__(kBlock) void k_spmv_csr(int N, const int *__restrict__ rptr,
                                                     const int *__restrict__ cind, const T *__restrict__ val,
                                                     const T *__restrict__ x, T *__restrict__ y,
                                                     const int *__restrict__ stop) {
    if (stop && *stop) return;
    const long long gid = (long long)xcd_block() * kBlock + threadIdx.x;
    const int row = (int)(gid / G), q = (int)(gid % G);
    T acc = 0;
    if (row < N) {
        const int e = rptr[row + 1];
        // chunks of 4 independent gathers per lane, summed in entry order
        for (int j0 = rptr[row] + q; j0 < e; j0 += 4 * G) {
            T xv[4], vv[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int j = j0 + u * G;
                if (j < e) {
                    vv[u] = val[j];
                    xv[u] = x[cind[j]];
                }
            }
#pragma unroll
            for (int u = 0; u < 4; u++)
                if (j0 + u * G < e) acc += vv[u] * xv[u];
        }
    }
    acc = group_sum<T, G>(acc);
    if (q == 0 && row < N) y[row] = acc;
}

template <typename T>
int launch_spmv_csr(Ctx *c, int N, const int *rptr, const int *cind, const T *val, const T *x, T *y,
                    double avg_row, const int *stop) {
    if (N <= 0) return SENK_OK;
    int G = avg_row > 48 ? 64 : avg_row > 24 ? 16 : avg_row > 12 ? 8 : avg_row > 6 ? 4 : avg_row > 3 ? 2 : 1;
    long long threads = (long long)N * G;
    int grid = (int)((threads + kBlock - 1) / kBlock);
    switch (G) {
        case 1: k_spmv_csr<T, 1><<<grid, kBlock, 0, c->stream>>>(N, rptr, cind, val, x, y, stop); break;
        case 2: k_spmv_csr<T, 2><<<grid, kBlock, 0, c->stream>>>(N, rptr, cind, val, x, y, stop); break;
        case 4: k_spmv_csr<T, 4><<<grid, kBlock, 0, c->stream>>>(N, rptr, cind, val, x, y, stop); break;
        case 8: k_spmv_csr<T, 8><<<grid, kBlock, 0, c->stream>>>(N, rptr, cind, val, x, y, stop); break;
        case 16: k_spmv_csr<T, 16><<<grid, kBlock, 0, c->stream>>>(N, rptr, cind, val, x, y, stop); break;
        default: k_spmv_csr<T, 64><<<grid, kBlock, 0, c->stream>>>(N, rptr, cind, val, x, y, stop); break;
    }
    SENK_LAUNCH_CHECK();
    return SENK_OK;
}
template int launch_spmv_csr<double>(Ctx *, int, const int *, const int *, const double *, const double *,
                                     double *, double, const int *);
template int launch_spmv_csr<float>(Ctx *, int, const int *, const int *, const float *, const float *, float *,
                                    double, const int *);

// senk_sparse.hpp:47-57 form: the diagonal product is the first addend.
__global__ __launch_bounds__(kBlock) void k_spmv_csr_diag(int N, const int *__restrict__ rptr,
                                                          const int *__restrict__ cind,
                                                          const double *__restrict__ val,
                                                          const double *__restrict__ diag,
                                                          const double *__restrict__ x, double *__restrict__ y) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= N) return;
    double acc = x[i] * diag[i];
    for (int j = rptr[i]; j < rptr[i + 1]; j++) acc += val[j] * x[cind[j]];
    y[i] = acc;
}
int launch_spmv_csr_diag_f64(Ctx *c, int N, const int *rptr, const int *cind, const double *val,
                             const double *diag, const double *x, double *y) {
    if (N <= 0) return SENK_OK;
    k_spmv_csr_diag<<<(N + kBlock - 1) / kBlock, kBlock, 0, c->stream>>>(N, rptr, cind, val, diag, x, y);
    SENK_LAUNCH_CHECK();
    return SENK_OK;
}

// =============================================================== SpMV SELL-C(-sigma)
// One lane per packed row; with C = 64 each slice is one wave and every slot load is
// a contiguous 64-element access.  Slice s has rows h = C (or N % C for a short last
// slice in the reference's Csr2Sell layout) and slot j of row k at
// wid[s]*C + j*h + k.  perm (optional) maps a packed row to its output row.
// O32: every element offset of the matrix stream and of x fits 32 bits in bytes, so each
// load is addressed as the array's (uniform) base plus one 32-bit register instead of a
// 64-bit address pair.  FP32: 56 instead of 70 VGPRs in chunks of 8 (8 waves per SIMD
// instead of 7; 128^3 7-point product 23.4 against 24.5-25.4 us), 96 instead of 126 in
// chunks of 16; FP64 keeps 90-92 / 159-162.
template <typename T>
__device__ __forceinline__ const T *at32(const T *p, unsigned i) {
    return reinterpret_cast<const T *>(reinterpret_cast<const char *>(p) + (size_t)(i * (unsigned)sizeof(T)));
}
template <typename T, int CH, bool NT, bool O32>
__global__ __launch_bounds__(kBlock) void k_spmv_sell(int N, int C, int nslices, const int *__restrict__ wid,
                                                      const int *__restrict__ cind, const T *__restrict__ val,
                                                      const int *__restrict__ perm, const T *__restrict__ x,
                                                      T *__restrict__ y, int last_short,
                                                      const int *__restrict__ stop,
                                                      const int *__restrict__ order) {
    if (stop && *stop) return;
    const long long gid = (long long)xcd_block() * kBlock + threadIdx.x;
    const int q = (int)(gid / C), k = (int)(gid % C);
    if (q >= nslices) return;
    const int s = order ? order[q] : q;
    const int h = (last_short && s == nslices - 1) ? N - s * C : C;
    if (k >= h) return;
    const size_t base = (size_t)wid[s] * C + k;
    const int w = wid[s + 1] - wid[s];
    T acc = 0;
    // slot 0 starts the sum (as the reference's SpmvSell), later slots in order; CH slots
    // of indices and values are loaded before their gathers are issued.  NT: the matrix
    // stream is read with non-temporal loads, which leave more of the L2 to x; whether that
    // pays depends on the matrix (measured per matrix, see launch_spmv_sell's callers).
    for (int j0 = 0; j0 < w; j0 += CH) {
        int cc[CH];
        T vv[CH], xv[CH];
#pragma unroll
        for (int u = 0; u < CH; u++) {
            const int j = j0 + u;
            if (j < w) {
                const size_t o = base + (size_t)j * h;
                const int *pc = O32 ? at32(cind, (unsigned)o) : cind + o;
                const T *pv = O32 ? at32(val, (unsigned)o) : val + o;
                if (NT) {
                    cc[u] = __builtin_nontemporal_load(pc);
                    vv[u] = __builtin_nontemporal_load(pv);
                } else {
                    cc[u] = *pc;
                    vv[u] = *pv;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < CH; u++)
            if (j0 + u < w) xv[u] = O32 ? *at32(x, (unsigned)cc[u]) : x[cc[u]];
#pragma unroll
        for (int u = 0; u < CH; u++) {
            const int j = j0 + u;
            if (j == 0) acc = vv[u] * xv[u];
            else if (j < w) acc += vv[u] * xv[u];
        }
    }
    const int row = perm ? perm[(size_t)s * C + k] : s * C + k;
    if (row >= 0) y[row] = acc;
}
template <typename T, bool NT, bool O32>
static void spmv_sell_launch(Ctx *c, int grid, bool short_rows, int N, int C, int nslices, const int *wid,
                             const int *cind, const T *val, const int *perm, const T *x, T *y, int last_short,
                             const int *stop, const int *order) {
    if (short_rows)
        k_spmv_sell<T, 8, NT, O32><<<grid, kBlock, 0, c->stream>>>(N, C, nslices, wid, cind, val, perm, x, y,
                                                                   last_short, stop, order);
    else
        k_spmv_sell<T, 16, NT, O32><<<grid, kBlock, 0, c->stream>>>(N, C, nslices, wid, cind, val, perm, x, y,
                                                                    last_short, stop, order);
}
template <typename T>
int launch_spmv_sell(Ctx *c, int N, int C, int nslices, const int *wid, const int *cind, const T *val,
                     const int *perm, const T *x, T *y, bool last_short, const int *stop, double avg_width,
                     const int *order, bool nontemporal, long long slots) {
    if (N <= 0 || nslices <= 0) return SENK_OK;
    long long threads = (long long)nslices * C;
    int grid = (int)((threads + kBlock - 1) / kBlock);
    // short rows (7-point stencils) keep all their slots in one chunk of 8; longer rows
    // take chunks of 16
    const bool short_rows = avg_width > 0 && avg_width <= 8.5;
    // 32-bit byte offsets where the stored slots (known) and x fit them
    const long long lim = (1LL << 32) / (long long)sizeof(T);
    const bool o32 = slots >= 0 && slots < lim && (long long)N < lim;
#define SENK_SELL(NTV, O32V)                                                                                  \
    spmv_sell_launch<T, NTV, O32V>(c, grid, short_rows, N, C, nslices, wid, cind, val, perm, x, y,            \
                                   last_short ? 1 : 0, stop, order)
    if (nontemporal && o32) SENK_SELL(true, true);
    else if (nontemporal) SENK_SELL(true, false);
    else if (o32) SENK_SELL(false, true);
    else SENK_SELL(false, false);
#undef SENK_SELL
    SENK_LAUNCH_CHECK();
    return SENK_OK;
}
template int launch_spmv_sell<double>(Ctx *, int, int, int, const int *, const int *, const double *, const int *,
                                      const double *, double *, bool, const int *, double, const int *, bool,
                                      long long);
template int launch_spmv_sell<float>(Ctx *, int, int, int, const int *, const int *, const float *, const int *,
                                     const float *, float *, bool, const int *, double, const int *, bool,
                                     long long);

// =============================================================== SpMV BCSR
// G lanes per block row; lane q takes blocks brptr[g]+q, +G, ...; each block is
// BNL*BNW contiguous values (column-major inside the block).  Partial row sums are
// combined inside the lane group.
template <typename T, int BNL, int BNW, int G>
__global__ __launch_bounds__(kBlock) void k_spmv_bcsr(int nbr, const int *__restrict__ brptr,
                                                      const int *__restrict__ bcind, const T *__restrict__ bval,
                                                      const T *__restrict__ x, T *__restrict__ y) {
    const long long gid = (long long)blockIdx.x * kBlock + threadIdx.x;
    const int g = (int)(gid / G), q = (int)(gid % G);
    T acc[BNL];
#pragma unroll
    for (int k = 0; k < BNL; k++) acc[k] = 0;
    if (g < nbr) {
        const int e = brptr[g + 1];
        for (int b = brptr[g] + q; b < e; b += G) {
            const T *blk = bval + (size_t)b * (BNL * BNW);
            const int xc = bcind[b] * BNW;
#pragma unroll
            for (int l = 0; l < BNW; l++) {
                const T xv = x[xc + l];
#pragma unroll
                for (int k = 0; k < BNL; k++) acc[k] += blk[l * BNL + k] * xv;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < BNL; k++) acc[k] = group_sum<T, G>(acc[k]);
    if (q == 0 && g < nbr) {
#pragma unroll
        for (int k = 0; k < BNL; k++) y[(size_t)g * BNL + k] = acc[k];
    }
}
// Any shape: one thread per block row, the reference's loop order.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_spmv_bcsr_any(int nbr, int bnl, int bnw, const int *__restrict__ brptr,
                                                          const int *__restrict__ bcind, const T *__restrict__ bval,
                                                          const T *__restrict__ x, T *__restrict__ y) {
    const int g = blockIdx.x * kBlock + threadIdx.x;
    if (g >= nbr) return;
    const int i = g * bnl;
    for (int k = 0; k < bnl; k++) y[i + k] = 0;
    for (int b = brptr[g]; b < brptr[g + 1]; b++) {
        const int xc = bcind[b] * bnw;
        for (int l = 0; l < bnw; l++)
            for (int k = 0; k < bnl; k++) y[i + k] += bval[(size_t)b * bnl * bnw + l * bnl + k] * x[xc + l];
    }
}
template <typename T>
int launch_spmv_bcsr(Ctx *c, int N, int bnl, int bnw, const int *brptr, const int *bcind, const T *bval,
                     const T *x, T *y) {
    if (N <= 0) return SENK_OK;
    const int nbr = N / bnl;
    auto grid_for = [&](int G) { return (int)(((long long)nbr * G + kBlock - 1) / kBlock); };
    if (bnl == 3 && bnw == 3) {
        k_spmv_bcsr<T, 3, 3, 4><<<grid_for(4), kBlock, 0, c->stream>>>(nbr, brptr, bcind, bval, x, y);
    } else if (bnl == 2 && bnw == 2) {
        k_spmv_bcsr<T, 2, 2, 4><<<grid_for(4), kBlock, 0, c->stream>>>(nbr, brptr, bcind, bval, x, y);
    } else if (bnl == 4 && bnw == 4) {
        k_spmv_bcsr<T, 4, 4, 4><<<grid_for(4), kBlock, 0, c->stream>>>(nbr, brptr, bcind, bval, x, y);
    } else if (bnl == 8 && bnw == 1) {
        k_spmv_bcsr<T, 8, 1, 8><<<grid_for(8), kBlock, 0, c->stream>>>(nbr, brptr, bcind, bval, x, y);
    } else if (bnl == 2 && bnw == 1) {
        k_spmv_bcsr<T, 2, 1, 8><<<grid_for(8), kBlock, 0, c->stream>>>(nbr, brptr, bcind, bval, x, y);
    } else if (bnl == 1 && bnw == 1) {
        k_spmv_bcsr<T, 1, 1, 8><<<grid_for(8), kBlock, 0, c->stream>>>(nbr, brptr, bcind, bval, x, y);
    } else {
        k_spmv_bcsr_any<T><<<(nbr + kBlock - 1) / kBlock, kBlock, 0, c->stream>>>(nbr, bnl, bnw, brptr, bcind, bval,
                                                                                  x, y);
    }
    SENK_LAUNCH_CHECK();
    return SENK_OK;
}
template int launch_spmv_bcsr<double>(Ctx *, int, int, int, const int *, const int *, const double *,
                                      const double *, double *);
template int launch_spmv_bcsr<float>(Ctx *, int, int, int, const int *, const int *, const float *, const float *,
                                     float *);

// =============================================================== SpTRSV (CSR)
// One launch per schedule level.  Packed form (BS | 64): one lane per row, the BS rows
// of a block are consecutive lanes of one wave.
//  phase 1: each lane subtracts the entries of its row that lie outside its block
//           (earlier levels, already final) in CSR order;
//  phase 2: BS steps; at step c row c of every block is final and is broadcast
//           inside the wave, and rows holding column base+c subtract it.
// Per row this is the reference's order (external columns precede internal ones for
// L, follow them for U, which walks descending), so results match the sequential
// solve bit for bit.  U rows carry the inverted diagonal first (Split "L-DU").
template <typename T, int BS, bool UPPER>
__global__ __launch_bounds__(kBlock) void k_sptrsv_packed(int nblk, int blk0, const int *__restrict__ blk_list,
                                                          const int *__restrict__ rptr,
                                                          const int *__restrict__ cind,
                                                          const T *__restrict__ val, const T *x, T *y,
                                                          const double *__restrict__ in_scale,
                                                          const int *__restrict__ stop_flag) {
    if (stop_flag && *stop_flag) return;
    const int gid = blockIdx.x * kBlock + threadIdx.x;
    const int bl = gid / BS, r = gid % BS;
    const bool active = bl < nblk;
    const int blk = active ? (blk_list ? blk_list[blk0 + bl] : blk0 + bl) : 0;
    const int base = blk * BS, row = base + r;
    T t = 0, dinv = 0;
    int jin = 0, jlim = 0;
    if (active) {
        const int s = rptr[row], e = rptr[row + 1];
        if (!UPPER) {
            t = x[row];
            if (in_scale) t = (T)((double)t * *in_scale);
            int j = s;
            for (; j < e; j++) {
                const int col = cind[j];
                if (col >= base) break;
                t -= val[j] * y[col];
            }
            jin = j;
            jlim = e;
        } else {
            t = x[row];
            dinv = val[s];
            int j = e - 1;
            for (; j > s; j--) {
                const int col = cind[j];
                if (col < base + BS) break;
                t -= val[j] * y[col];
            }
            jin = j;
            jlim = s;
        }
    }
    if (BS > 1) {
        const int lane = threadIdx.x & 63, lbase = lane - r;
        if (!UPPER) {
#pragma unroll 4
            for (int c = 0; c < BS; c++) {
                const T yc = __shfl(t, lbase + c, 64);
                if (active && jin < jlim && cind[jin] == base + c) {
                    t -= val[jin] * yc;
                    jin++;
                }
            }
        } else {
#pragma unroll 4
            for (int c = BS - 1; c >= 0; c--) {
                if (r == c) t = t * dinv;
                const T yc = __shfl(t, lbase + c, 64);
                if (active && jin > jlim && cind[jin] == base + c) {
                    t -= val[jin] * yc;
                    jin--;
                }
            }
        }
    } else if (UPPER) {
        t = t * dinv;
    }
    if (active) y[row] = t;
}

// LDS-staged form for launches over a contiguous row range (colors of an ABMC
// ordering are contiguous): a workgroup of kBlock consecutive rows first copies its
// rows' CSR slice [rptr[r0], rptr[r0+kBlock]) into LDS with coalesced loads, then runs
// the same two phases on the staged copy.  Per row the arithmetic and its order are
// those of k_sptrsv_packed (bit-identical results); only the memory path changes.
// A slice larger than the LDS buffer falls back to reading global memory.
template <typename T, int BS, bool UPPER, bool STAGED>
__device__ __forceinline__ void sptrsv_rows(int row, bool active, const int *__restrict__ rptr,
                                            const int *__restrict__ cind, const T *__restrict__ val, const int *lc,
                                            const T *lv, int e0, const T *x, T *y, const double *in_scale) {
    auto C = [&](int j) { return STAGED ? lc[j - e0] : cind[j]; };
    auto V = [&](int j) { return STAGED ? lv[j - e0] : val[j]; };
    constexpr int CH = 8;  // gathers issued together
    const int r = row % BS, base = row - r;
    T t = 0, dinv = 0;
    int jin = 0, jlim = 0;
    if (active) {
        const int s = rptr[row], e = rptr[row + 1];
        t = x[row];
        if (!UPPER) {
            if (in_scale) t = (T)((double)t * *in_scale);
            // external entries are the prefix with col < base; find its end first so the
            // gathers below do not wait on each other
            int je = s;
            while (je < e && C(je) < base) je++;
            for (int j0 = s; j0 < je; j0 += CH) {
                T yv[CH], vv[CH];
#pragma unroll
                for (int q = 0; q < CH; q++)
                    if (j0 + q < je) {
                        vv[q] = V(j0 + q);
                        yv[q] = y[C(j0 + q)];
                    }
#pragma unroll
                for (int q = 0; q < CH; q++)
                    if (j0 + q < je) t -= vv[q] * yv[q];
            }
            jin = je;
            jlim = e;
        } else {
            dinv = V(s);
            // external entries are the suffix with col >= base + BS, consumed descending
            int je = e - 1;
            while (je > s && C(je) >= base + BS) je--;
            for (int j0 = e - 1; j0 > je; j0 -= CH) {
                T yv[CH], vv[CH];
#pragma unroll
                for (int q = 0; q < CH; q++)
                    if (j0 - q > je) {
                        vv[q] = V(j0 - q);
                        yv[q] = y[C(j0 - q)];
                    }
#pragma unroll
                for (int q = 0; q < CH; q++)
                    if (j0 - q > je) t -= vv[q] * yv[q];
            }
            jin = je;
            jlim = s;
        }
    }
    if (BS > 1) {
        const int lane = threadIdx.x & 63, lbase = lane - r;
        if (!UPPER) {
#pragma unroll 4
            for (int c = 0; c < BS; c++) {
                const T yc = __shfl(t, lbase + c, 64);
                if (active && jin < jlim && C(jin) == base + c) {
                    t -= V(jin) * yc;
                    jin++;
                }
            }
        } else {
#pragma unroll 4
            for (int c = BS - 1; c >= 0; c--) {
                if (r == c) t = t * dinv;
                const T yc = __shfl(t, lbase + c, 64);
                if (active && jin > jlim && C(jin) == base + c) {
                    t -= V(jin) * yc;
                    jin--;
                }
            }
        }
    } else if (UPPER) {
        t = t * dinv;
    }
    if (active) y[row] = t;
}

template <typename T, int BS, bool UPPER>
__global__ __launch_bounds__(kBlock) void k_sptrsv_lds(int row0, int nrows, const int *__restrict__ rptr,
                                                       const int *__restrict__ cind, const T *__restrict__ val,
                                                       const T *x, T *y, const double *__restrict__ in_scale,
                                                       const int *__restrict__ stop_flag) {
    if (stop_flag && *stop_flag) return;
    constexpr int CAP = (40 * 1024) / (sizeof(T) + sizeof(int));
    __shared__ T lv[CAP];
    __shared__ int lc[CAP];
    const int r_begin = row0 + blockIdx.x * kBlock;
    const int r_end = min(row0 + nrows, r_begin + kBlock);
    const int e0 = rptr[r_begin], e1 = rptr[r_end];
    const int row = r_begin + threadIdx.x;
    const bool active = row < r_end;
    if (e1 - e0 <= CAP) {
        for (int q = threadIdx.x; q < e1 - e0; q += kBlock) {
            lv[q] = val[e0 + q];
            lc[q] = cind[e0 + q];
        }
        __syncthreads();
        sptrsv_rows<T, BS, UPPER, true>(row, active, rptr, cind, val, lc, lv, e0, x, y, in_scale);
    } else {
        sptrsv_rows<T, BS, UPPER, false>(row, active, rptr, cind, val, lc, lv, e0, x, y, in_scale);
    }
}

// Any block size: one wave per block, rows in order, the wave splits each row's
// entries and reduces them; in-block values come from LDS.  (Summation order differs
// from the reference's, within rounding.)
template <typename T, bool UPPER>
__global__ __launch_bounds__(64) void k_sptrsv_wave(int nblk, int blk0, const int *__restrict__ blk_list, int BS,
                                                    const int *__restrict__ rptr, const int *__restrict__ cind,
                                                    const T *__restrict__ val, const T *x, T *y,
                                                    const double *__restrict__ in_scale,
                                                    const int *__restrict__ stop_flag) {
    extern __shared__ double ybuf_raw[];
    T *ybuf = reinterpret_cast<T *>(ybuf_raw);
    if (stop_flag && *stop_flag) return;
    const int bl = blockIdx.x;
    if (bl >= nblk) return;
    const int blk = blk_list ? blk_list[blk0 + bl] : blk0 + bl;
    const int base = blk * BS, lane = threadIdx.x;
    for (int step = 0; step < BS; step++) {
        const int r = UPPER ? BS - 1 - step : step;
        const int row = base + r;
        const int s = rptr[row], e = rptr[row + 1];
        double part = 0;
        for (int j = (UPPER ? s + 1 : s) + lane; j < e; j += 64) {
            const int col = cind[j];
            const T yc = (col >= base && col < base + BS) ? ybuf[col - base] : y[col];
            part += (double)(val[j] * yc);
        }
        const double tot = wave_sum(part);
        if (lane == 0) {
            T xv = x[row];
            if (!UPPER && in_scale) xv = (T)((double)xv * *in_scale);
            T res = UPPER ? (T)((double)xv - tot) * val[s] : (T)((double)xv - tot);
            ybuf[r] = res;
            y[row] = res;
        }
        __syncthreads();
    }
}

template <typename T>
int launch_sptrsv(Ctx *c, const Schedule &s, bool upper, const int *rptr, const int *cind, const T *val,
                  const T *x, T *y, const double *in_scale, const int *stop_flag) {
    const int BS = s.bsize;
    const bool packed = BS <= 64 && (64 % BS) == 0;
    for (int step = 0; step < s.nlev; step++) {
        const int l = upper ? s.nlev - 1 - step : step;
        const int b0 = s.lev_ptr[l], nb = s.lev_ptr[l + 1] - b0;
        if (nb <= 0) continue;
        // an upper solve reads x of rows not yet solved only in its own row, so the
        // scale applies to the lower solve only
        if (packed && !s.dblk) {
            // contiguous rows [b0*BS, (b0+nb)*BS): LDS-staged kernel
            const int row0 = b0 * BS, nrows = nb * BS;
            const int grid = (nrows + kBlock - 1) / kBlock;
#define SENK_TRSV_LDS(B)                                                                                        \
    case B:                                                                                                     \
        if (upper)                                                                                              \
            k_sptrsv_lds<T, B, true><<<grid, kBlock, 0, c->stream>>>(row0, nrows, rptr, cind, val, x, y, nullptr, \
                                                                     stop_flag);                                \
        else                                                                                                    \
            k_sptrsv_lds<T, B, false><<<grid, kBlock, 0, c->stream>>>(row0, nrows, rptr, cind, val, x, y,         \
                                                                      in_scale, stop_flag);                     \
        break;
            switch (BS) {
                SENK_TRSV_LDS(1)
                SENK_TRSV_LDS(2)
                SENK_TRSV_LDS(4)
                SENK_TRSV_LDS(8)
                SENK_TRSV_LDS(16)
                SENK_TRSV_LDS(32)
                SENK_TRSV_LDS(64)
                default: return SENK_ERR_UNSUPPORTED;
            }
#undef SENK_TRSV_LDS
        } else if (packed) {
            const long long threads = (long long)nb * BS;
            const int grid = (int)((threads + kBlock - 1) / kBlock);
#define SENK_TRSV_CASE(B)                                                                                        \
    case B:                                                                                                      \
        if (upper)                                                                                               \
            k_sptrsv_packed<T, B, true><<<grid, kBlock, 0, c->stream>>>(nb, b0, s.dblk, rptr, cind, val, x, y,   \
                                                                        nullptr, stop_flag);                     \
        else                                                                                                     \
            k_sptrsv_packed<T, B, false><<<grid, kBlock, 0, c->stream>>>(nb, b0, s.dblk, rptr, cind, val, x, y,  \
                                                                         in_scale, stop_flag);                   \
        break;
            switch (BS) {
                SENK_TRSV_CASE(1)
                SENK_TRSV_CASE(2)
                SENK_TRSV_CASE(4)
                SENK_TRSV_CASE(8)
                SENK_TRSV_CASE(16)
                SENK_TRSV_CASE(32)
                SENK_TRSV_CASE(64)
                default: return SENK_ERR_UNSUPPORTED;
            }
#undef SENK_TRSV_CASE
        } else {
            const size_t lds = sizeof(T) * (size_t)BS;
            if (lds > 64 * 1024) return SENK_ERR_UNSUPPORTED;
            if (upper)
                k_sptrsv_wave<T, true><<<nb, 64, lds, c->stream>>>(nb, b0, s.dblk, BS, rptr, cind, val, x, y,
                                                                   nullptr, stop_flag);
            else
                k_sptrsv_wave<T, false><<<nb, 64, lds, c->stream>>>(nb, b0, s.dblk, BS, rptr, cind, val, x, y,
                                                                    in_scale, stop_flag);
        }
        SENK_LAUNCH_CHECK();
    }
    return SENK_OK;
}
template int launch_sptrsv<double>(Ctx *, const Schedule &, bool, const int *, const int *, const double *,
                                   const double *, double *, const double *, const int *);
template int launch_sptrsv<float>(Ctx *, const Schedule &, bool, const int *, const int *, const float *,
                                  const float *, float *, const double *, const int *);

// =============================================================== SpTRSV (BCSR)
// One lane per schedule block of bsize rows, walking its bnl-row groups in order with
// the reference's column-oriented updates (senk_sparse.hpp:399-415, 436-464); the
// group's values stay in registers.
template <typename T, int BNL>
__global__ __launch_bounds__(kBlock) void k_sptrsv_bcsr(int nblk, int blk0, const int *__restrict__ blk_list,
                                                        int bsize, int bnw, bool upper,
                                                        const int *__restrict__ brptr,
                                                        const int *__restrict__ bcind,
                                                        const T *__restrict__ bval, const T *x, T *y) {
    const int bl = blockIdx.x * kBlock + threadIdx.x;
    if (bl >= nblk) return;
    const int blk = blk_list ? blk_list[blk0 + bl] : blk0 + bl;
    const int bs = BNL * bnw, b_rem = BNL / bnw, ngroups = bsize / BNL;
    for (int gq = 0; gq < ngroups; gq++) {
        const int gi = upper ? ngroups - 1 - gq : gq;
        const int i = blk * bsize + gi * BNL, g = i / BNL;
        T yv[BNL];
#pragma unroll
        for (int k = 0; k < BNL; k++) yv[k] = x[i + k];
        if (!upper) {
            for (int b = brptr[g]; b < brptr[g + 1]; b++) {
                const int xc = bcind[b] * bnw;
                for (int l = 0; l < bnw; l++) {
                    const int col = xc + l;
                    T yc;
                    if (col >= i && col < i + BNL) {
                        yc = 0;
#pragma unroll
                        for (int k = 0; k < BNL; k++)
                            if (k == col - i) yc = yv[k];
                    } else {
                        yc = y[col];
                    }
                    const T *cb = bval + (size_t)b * bs + (size_t)l * BNL;
#pragma unroll
                    for (int k = 0; k < BNL; k++) yv[k] -= cb[k] * yc;
                }
            }
        } else {
            for (int b = brptr[g + 1] - 1; b >= brptr[g] + b_rem; b--) {
                const int xc = bcind[b] * bnw;
                for (int l = 0; l < bnw; l++) {
                    const T yc = y[xc + l];
                    const T *cb = bval + (size_t)b * bs + (size_t)l * BNL;
#pragma unroll
                    for (int k = 0; k < BNL; k++) yv[k] -= cb[k] * yc;
                }
            }
            int pos = brptr[g] + b_rem - 1;
            for (int qq = b_rem - 1; qq >= 0; qq--, pos--) {
                for (int l = bnw - 1; l >= 0; l--) {
                    const T *cb = bval + (size_t)pos * bs + (size_t)l * BNL;
                    const int d = qq * bnw + l;
                    T yd = 0;
#pragma unroll
                    for (int k = 0; k < BNL; k++)
                        if (k == d) { yv[k] *= cb[d]; yd = yv[k]; }
#pragma unroll
                    for (int k = 0; k < BNL; k++)
                        if (k < d) yv[k] -= cb[k] * yd;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < BNL; k++) y[i + k] = yv[k];
    }
}
template <typename T>
int launch_sptrsv_bcsr(Ctx *c, const Schedule &s, bool upper, int bnl, int bnw, const int *brptr, const int *bcind,
                       const T *bval, const T *x, T *y) {
    if (s.bsize % bnl) return SENK_ERR_DIVISIBILITY;
    for (int step = 0; step < s.nlev; step++) {
        const int l = upper ? s.nlev - 1 - step : step;
        const int b0 = s.lev_ptr[l], nb = s.lev_ptr[l + 1] - b0;
        if (nb <= 0) continue;
        const int grid = (nb + kBlock - 1) / kBlock;
        switch (bnl) {
            case 1: k_sptrsv_bcsr<T, 1><<<grid, kBlock, 0, c->stream>>>(nb, b0, s.dblk, s.bsize, bnw, upper, brptr, bcind, bval, x, y); break;
            case 2: k_sptrsv_bcsr<T, 2><<<grid, kBlock, 0, c->stream>>>(nb, b0, s.dblk, s.bsize, bnw, upper, brptr, bcind, bval, x, y); break;
            case 3: k_sptrsv_bcsr<T, 3><<<grid, kBlock, 0, c->stream>>>(nb, b0, s.dblk, s.bsize, bnw, upper, brptr, bcind, bval, x, y); break;
            case 4: k_sptrsv_bcsr<T, 4><<<grid, kBlock, 0, c->stream>>>(nb, b0, s.dblk, s.bsize, bnw, upper, brptr, bcind, bval, x, y); break;
            case 8: k_sptrsv_bcsr<T, 8><<<grid, kBlock, 0, c->stream>>>(nb, b0, s.dblk, s.bsize, bnw, upper, brptr, bcind, bval, x, y); break;
            default: return SENK_ERR_UNSUPPORTED;
        }
        SENK_LAUNCH_CHECK();
    }
    return SENK_OK;
}
template int launch_sptrsv_bcsr<double>(Ctx *, const Schedule &, bool, int, int, const int *, const int *,
                                        const double *, const double *, double *);
template int launch_sptrsv_bcsr<float>(Ctx *, const Schedule &, bool, int, int, const int *, const int *,
                                       const float *, const float *, float *);

// =============================================================== ILU(0) numeric
// One wave per schedule block; the block's rows are factored in order.  A row of up
// to 64*R entries lives in registers (lane q holds entries q, q+64, ...; R = 4, 16 or 64
// by the longest row of the factor, so rows of up to 4096 entries).  For every
// strictly-lower entry k (ascending): the owner lane divides it by the pivot row's
// diagonal, the quotient is broadcast, and every later entry of the row subtracts
// quotient * (pivot row entry of the same column), found by binary search in the pivot
// row's upper part.  Entries thus receive their updates in increasing pivot column,
// the reference's IKJ order (senk_matrix.hpp:418-448).  Pivot rows are final: earlier
// levels were written by earlier launches, earlier rows of the same block by this
// wave followed by a workgroup-scope fence.
template <typename T, int kIluRegs>
__global__ __launch_bounds__(kBlock) void k_ilu0(int nblk, int blk0, const int *__restrict__ blk_list, int BS,
                                                 const int *__restrict__ rptr, const int *__restrict__ cind,
                                                 const int *__restrict__ dpos, T *val, int *err) {
    const int wave = (blockIdx.x * kBlock + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    if (wave >= nblk) return;
    const int blk = blk_list ? blk_list[blk0 + wave] : blk0 + wave;
    for (int r = 0; r < BS; r++) {
        const int row = blk * BS + r;
        const int s = rptr[row], e = rptr[row + 1], len = e - s;
        if (len > 64 * kIluRegs) {
            if (lane == 0) atomicExch(err, 2);
            return;
        }
        T v[kIluRegs];
        int col[kIluRegs];
#pragma unroll
        for (int q = 0; q < kIluRegs; q++) {
            const int j = s + lane + 64 * q;
            v[q] = j < e ? val[j] : (T)0;
            col[q] = j < e ? cind[j] : 0x7fffffff;
        }
        for (int k = 0; k < len; k++) {
            const int ck = cind[s + k];   // uniform
            if (ck >= row) break;
            const int owner_q = k >> 6, owner_l = k & 63;
            // pivot value after all previous updates, held by its owner lane
            T piv = 0;
#pragma unroll
            for (int q = 0; q < kIluRegs; q++)
                if (q == owner_q) piv = __shfl(v[q], owner_l, 64);
            const int dp = dpos[ck];
            if (dp >= 0) {
                const T d = val[dp];
                if (d == (T)0) {
                    if (lane == 0) atomicExch(err, 1);
                } else {
                    piv = piv / d;
                }
            }
#pragma unroll
            for (int q = 0; q < kIluRegs; q++)
                if (q == owner_q && lane == owner_l) v[q] = piv;
            // updates of the later entries of the row
            const int ps = (dp >= 0 ? dp : rptr[ck]), pe = rptr[ck + 1];
#pragma unroll
            for (int q = 0; q < kIluRegs; q++) {
                const int jj = lane + 64 * q;
                if (jj > k && jj < len) {
                    const int cj = col[q];
                    int lo = ps, hi = pe - 1;
                    while (lo <= hi) {
                        const int mid = (lo + hi) >> 1;
                        const int cm = cind[mid];
                        if (cm == cj) {
                            v[q] -= piv * val[mid];
                            break;
                        }
                        if (cm < cj) lo = mid + 1;
                        else hi = mid - 1;
                    }
                }
            }
        }
#pragma unroll
        for (int q = 0; q < kIluRegs; q++) {
            const int j = s + lane + 64 * q;
            if (j < e) val[j] = v[q];
        }
        // the next rows of the block read this one from other lanes of the same wave: a
        // workgroup-scope fence orders it (later launches see it through the kernel
        // boundary; a device-scope fence here would write back the XCD's whole L2)
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    }
}
// Longest row: a grid-stride pass and one atomic per workgroup (one per wave over 2^21
// rows took 374 us, the atomics on one word serialised).
constexpr int kMaxRowLenGrid = 512;
__global__ __launch_bounds__(kBlock) void k_max_row_len(int N, const int *__restrict__ rptr, int *out) {
    __shared__ int sh[kWavesPerBlock];
    int len = 0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < N; i += gridDim.x * kBlock)
        len = max(len, rptr[i + 1] - rptr[i]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) len = max(len, __shfl_xor(len, off, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = len;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kWavesPerBlock; w++) len = max(len, sh[w]);
        if (len > 0) atomicMax(out, len);
    }
}
template <typename T, int R>
static int ilu0_levels(Ctx *c, const Schedule &s, const int *rptr, const int *cind, const int *dpos, T *val,
                       int *derr) {
    for (int l = 0; l < s.nlev; l++) {
        const int b0 = s.lev_ptr[l], nb = s.lev_ptr[l + 1] - b0;
        if (nb <= 0) continue;
        const int grid = (nb + kWavesPerBlock - 1) / kWavesPerBlock;
        k_ilu0<T, R><<<grid, kBlock, 0, c->stream>>>(nb, b0, s.dblk, s.bsize, rptr, cind, dpos, val, derr);
        SENK_LAUNCH_CHECK();
    }
    return SENK_OK;
}
template <typename T>
int launch_ilu0(Ctx *c, const Schedule &s, int N, const int *rptr, const int *cind, const int *dpos, T *val,
                int *derr) {
    if (N <= 0) return SENK_OK;
    // the longest row picks the register count (a row beyond 4096 entries: err = 2)
    int *dmax = nullptr, maxlen = 0;
    SENK_HIP(hipMalloc(&dmax, sizeof(int)));
    int st = SENK_OK;
    if (hipMemsetAsync(dmax, 0, sizeof(int), c->stream) != hipSuccess) st = SENK_ERR_HIP;
    if (!st) {
        k_max_row_len<<<std::max(1, std::min((N + kBlock - 1) / kBlock, kMaxRowLenGrid)), kBlock, 0, c->stream>>>(
            N, rptr, dmax);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(&maxlen, dmax, sizeof(int), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            st = SENK_ERR_HIP;
    }
    (void)hipFree(dmax);
    if (st) return st;
    if (maxlen <= 64 * 4) return ilu0_levels<T, 4>(c, s, rptr, cind, dpos, val, derr);
    if (maxlen <= 64 * 16) return ilu0_levels<T, 16>(c, s, rptr, cind, dpos, val, derr);
    return ilu0_levels<T, 64>(c, s, rptr, cind, dpos, val, derr);
}
template int launch_ilu0<double>(Ctx *, const Schedule &, int, const int *, const int *, const int *, double *,
                                 int *);
template int launch_ilu0<float>(Ctx *, const Schedule &, int, const int *, const int *, const int *, float *,
                                int *);

__global__ __launch_bounds__(kBlock) void k_find_diag(int N, const int *__restrict__ rptr,
                                                      const int *__restrict__ cind, int *__restrict__ dpos) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= N) return;
    int p = -1;
    for (int j = rptr[i]; j < rptr[i + 1]; j++)
        if (cind[j] == i) { p = j; break; }
    dpos[i] = p;
}
int launch_find_diag(Ctx *c, int N, const int *rptr, const int *cind, int *dpos) {
    if (N <= 0) return SENK_OK;
    k_find_diag<<<(N + kBlock - 1) / kBlock, kBlock, 0, c->stream>>>(N, rptr, cind, dpos);
    SENK_LAUNCH_CHECK();
    return SENK_OK;
}

// D-ILU vectors: d = diag(U) from U's inverted diagonal (first entry of each row),
// e = 2 d - diag(A), both rounded to FP32
__global__ __launch_bounds__(kBlock) void k_dilu_diag(int N, const int *__restrict__ urptr,
                                                      const double *__restrict__ uval, const int *__restrict__ arptr,
                                                      const int *__restrict__ acind, const double *__restrict__ aval,
                                                      float *d, float *e) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= N) return;
    const double di = 1.0 / uval[urptr[i]];
    double a = 0;
    for (int j = arptr[i]; j < arptr[i + 1]; j++)
        if (acind[j] == i) a = aval[j];
    d[i] = (float)di;
    e[i] = (float)(2.0 * di - a);
}
int launch_dilu_diag(Ctx *c, int N, const int *urptr, const double *uval, const int *arptr, const int *acind,
                     const double *aval, float *d, float *e) {
    if (N <= 0) return SENK_OK;
    k_dilu_diag<<<(N + kBlock - 1) / kBlock, kBlock, 0, c->stream>>>(N, urptr, uval, arptr, acind, aval, d, e);
    SENK_LAUNCH_CHECK();
    return SENK_OK;
}

// IC(0)/IC(p): the upper factor from the lower one, u_ij = d_i l_ji, so that L D L^T is
// exactly symmetric.  L entry e (row j, column i = lcol[e]) goes to U position umap[e];
// d_i is the factorization's pivot, f[dpos[i]] (the full combined factor).
__global__ __launch_bounds__(kBlock) void k_ic_upper(int lnnz, const int *__restrict__ umap,
                                                     const int *__restrict__ lcol, const double *__restrict__ f,
                                                     const int *__restrict__ dpos, const double *__restrict__ l,
                                                     double *u) {
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= lnnz) return;
    u[umap[e]] = f[dpos[lcol[e]]] * l[e];
}
int launch_ic_upper(Ctx *c, int lnnz, const int *umap, const int *lcol, const double *f, const int *dpos,
                    const double *l, double *u) {
    if (lnnz <= 0) return SENK_OK;
    k_ic_upper<<<(lnnz + kBlock - 1) / kBlock, kBlock, 0, c->stream>>>(lnnz, umap, lcol, f, dpos, l, u);
    SENK_LAUNCH_CHECK();
    return SENK_OK;
}

// Split "L-DU" with inverted diagonal (senk_matrix.hpp:257-333) on the device; the
// output row pointers come from the host pattern split.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_split_ldu(int N, const int *__restrict__ rptr,
                                                      const int *__restrict__ cind, const T *__restrict__ val,
                                                      const int *__restrict__ lrptr, T *__restrict__ lval,
                                                      int *__restrict__ lcind, const int *__restrict__ urptr,
                                                      T *__restrict__ uval, int *__restrict__ ucind) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= N) return;
    int nl = lrptr[i], nu = urptr[i];
    for (int j = rptr[i]; j < rptr[i + 1]; j++) {
        const int col = cind[j];
        if (col < i) { lval[nl] = val[j]; lcind[nl] = col; nl++; }
        else if (col > i) { uval[nu] = val[j]; ucind[nu] = col; nu++; }
        else { uval[nu] = (T)1 / val[j]; ucind[nu] = col; nu++; }
    }
}
template <typename T>
int launch_split_ldu(Ctx *c, int N, const int *rptr, const int *cind, const T *val, const int *lrptr, T *lval,
                     int *lcind, const int *urptr, T *uval, int *ucind) {
    if (N <= 0) return SENK_OK;
    k_split_ldu<T><<<(N + kBlock - 1) / kBlock, kBlock, 0, c->stream>>>(N, rptr, cind, val, lrptr, lval, lcind, urptr,
                                                                        uval, ucind);
    SENK_LAUNCH_CHECK();
    return SENK_OK;
}
template int launch_split_ldu<double>(Ctx *, int, const int *, const int *, const double *, const int *, double *,
                                      int *, const int *, double *, int *);
template int launch_split_ldu<float>(Ctx *, int, const int *, const int *, const float *, const int *, float *,
                                     int *, const int *, float *, int *);

// =============================================================== BLAS-1
inline int ew_grid(long long n) {
    long long b = (n + kBlock * 4LL - 1) / (kBlock * 4LL);
    if (b < 1) b = 1;
    if (b > 8192) b = 8192;
    return (int)b;
}
#define SENK_GRID_STRIDE(i, n) \
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < (n); i += (long long)gridDim.x * kBlock)

// Elementwise passes: four consecutive elements per thread, 16-byte accesses.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_axpby(int N, T a, const T *__restrict__ x, T b, T *y) {
    SENK_FOR4(i0, N) {
        T xv[4], yv[4];
        get4(x, i0, N, xv);
        get4(y, i0, N, yv);
#pragma unroll
        for (int u = 0; u < 4; u++) yv[u] = a * xv[u] + b * yv[u];
        put4(y, i0, N, yv);
    }
}
template <typename T>
__global__ __launch_bounds__(kBlock) void k_axpy(int N, T a, const T *__restrict__ x, T *y) {
    SENK_FOR4(i0, N) {
        T xv[4], yv[4];
        get4(x, i0, N, xv);
        get4(y, i0, N, yv);
#pragma unroll
        for (int u = 0; u < 4; u++) yv[u] += a * xv[u];
        put4(y, i0, N, yv);
    }
}
template <typename T>
__global__ __launch_bounds__(kBlock) void k_axpyz(int N, T a, const T *x, const T *y, T *z) {
    SENK_FOR4(i0, N) {
        T xv[4], yv[4];
        get4(x, i0, N, xv);
        get4(y, i0, N, yv);
#pragma unroll
        for (int u = 0; u < 4; u++) yv[u] = a * xv[u] + yv[u];
        put4(z, i0, N, yv);
    }
}
template <typename T>
__global__ __launch_bounds__(kBlock) void k_scal(int N, T a, T *x) {
    SENK_FOR4(i0, N) {
        T xv[4];
        get4(x, i0, N, xv);
#pragma unroll
        for (int u = 0; u < 4; u++) xv[u] *= a;
        put4(x, i0, N, xv);
    }
}
template <typename T>
__global__ __launch_bounds__(kBlock) void k_copy(int N, const T *__restrict__ x, T *__restrict__ y) {
    SENK_FOR4(i0, N) {
        T xv[4];
        get4(x, i0, N, xv);
        put4(y, i0, N, xv);
    }
}
template <typename T>
__global__ __launch_bounds__(kBlock) void k_hadprod(int N, const T *__restrict__ x, T *y) {
    SENK_FOR4(i0, N) {
        T xv[4], yv[4];
        get4(x, i0, N, xv);
        get4(y, i0, N, yv);
#pragma unroll
        for (int u = 0; u < 4; u++) yv[u] *= xv[u];
        put4(y, i0, N, yv);
    }
}
template <typename T>
__global__ __launch_bounds__(kBlock) void k_haddiv(int N, const T *__restrict__ x, T *y) {
    SENK_FOR4(i0, N) {
        T xv[4], yv[4];
        get4(x, i0, N, xv);
        get4(y, i0, N, yv);
#pragma unroll
        for (int u = 0; u < 4; u++) yv[u] /= xv[u];
        put4(y, i0, N, yv);
    }
}
template <typename TI, typename TO>
__global__ __launch_bounds__(kBlock) void k_convert(int N, const TI *__restrict__ in, TO *__restrict__ out,
                                                    const double *__restrict__ dscale) {
    const double sc = dscale ? *dscale : 1.0;
    SENK_FOR4(i0, N) {
        TI a[4];
        TO b[4];
        get4(in, i0, N, a);
        if (dscale) {
#pragma unroll
            for (int u = 0; u < 4; u++) b[u] = (TO)((double)a[u] * sc);
        } else {
#pragma unroll
            for (int u = 0; u < 4; u++) b[u] = (TO)a[u];
        }
        put4(out, i0, N, b);
    }
}

template <typename T>
int launch_axpby(Ctx *c, int N, double a, const T *x, double b, T *y) {
    if (N <= 0) return SENK_OK;
    if (b == 1.0)
        k_axpy<T><<<ew_grid(N), kBlock, 0, c->stream>>>(N, (T)a, x, y);
    else
        k_axpby<T><<<ew_grid(N), kBlock, 0, c->stream>>>(N, (T)a, x, (T)b, y);
    SENK_LAUNCH_CHECK();
    return SENK_OK;
}
template <typename T>
int launch_axpyz(Ctx *c, int N, double a, const T *x, const T *y, T *z) {
    if (N <= 0) return SENK_OK;
    k_axpyz<T><<<ew_grid(N), kBlock, 0, c->stream>>>(N, (T)a, x, y, z);
    SENK_LAUNCH_CHECK();
    return SENK_OK;
}
template <typename T>
int launch_scal(Ctx *c, int N, double a, T *x) {
    if (N <= 0) return SENK_OK;
    k_scal<T><<<ew_grid(N), kBlock, 0, c->stream>>>(N, (T)a, x);
    SENK_LAUNCH_CHECK();
    return SENK_OK;
}
template <typename T>
int launch_copy(Ctx *c, int N, const T *x, T *y) {
    if (N <= 0) return SENK_OK;
    k_copy<T><<<ew_grid(N), kBlock, 0, c->stream>>>(N, x, y);
    SENK_LAUNCH_CHECK();
    return SENK_OK;
}
template <typename T>
int launch_hadprod(Ctx *c, int N, const T *x, T *y) {
    if (N <= 0) return SENK_OK;
    k_hadprod<T><<<ew_grid(N), kBlock, 0, c->stream>>>(N, x, y);
    SENK_LAUNCH_CHECK();
    return SENK_OK;
}
template <typename T>
int launch_haddiv(Ctx *c, int N, const T *x, T *y) {
    if (N <= 0) return SENK_OK;
    k_haddiv<T><<<ew_grid(N), kBlock, 0, c->stream>>>(N, x, y);
    SENK_LAUNCH_CHECK();
    return SENK_OK;
}
template <typename TI, typename TO>
int launch_convert(Ctx *c, int N, const TI *in, TO *out, const double *dscale) {
    if (N <= 0) return SENK_OK;
    k_convert<TI, TO><<<ew_grid(N), kBlock, 0, c->stream>>>(N, in, out, dscale);
    SENK_LAUNCH_CHECK();
    return SENK_OK;
}
template int launch_axpby<double>(Ctx *, int, double, const double *, double, double *);
template int launch_axpby<float>(Ctx *, int, double, const float *, double, float *);
template int launch_axpyz<double>(Ctx *, int, double, const double *, const double *, double *);
template int launch_axpyz<float>(Ctx *, int, double, const float *, const float *, float *);
template int launch_scal<double>(Ctx *, int, double, double *);
template int launch_scal<float>(Ctx *, int, double, float *);
template int launch_copy<double>(Ctx *, int, const double *, double *);
template int launch_copy<float>(Ctx *, int, const float *, float *);
template int launch_hadprod<double>(Ctx *, int, const double *, double *);
template int launch_hadprod<float>(Ctx *, int, const float *, float *);
template int launch_haddiv<double>(Ctx *, int, const double *, double *);
template int launch_haddiv<float>(Ctx *, int, const float *, float *);
template int launch_convert<double, float>(Ctx *, int, const double *, float *, const double *);
template int launch_convert<float, double>(Ctx *, int, const float *, double *, const double *);
template int launch_convert<double, double>(Ctx *, int, const double *, double *, const double *);

// K simultaneous dot products against one vector, FP64 accumulation: a streaming
// pass writes per-workgroup partials, then one workgroup per product adds them in a
// fixed order.
template <typename T, int KMAX>
__global__ __launch_bounds__(kBlock) void k_multidot(int N, int K, const T *__restrict__ xs, size_t ldx,
                                                     const T *__restrict__ y, double *partials) {
    double acc[KMAX];
#pragma unroll
    for (int k = 0; k < KMAX; k++) acc[k] = 0;
    SENK_FOR4(i0, N) {
        T yv[4];
        get4(y, i0, N, yv);
#pragma unroll
        for (int k = 0; k < KMAX; k++)
            if (k < K) {
                T xv[4];
                get4(xs + (size_t)k * ldx, i0, N, xv);
#pragma unroll
                for (int u = 0; u < 4; u++) acc[k] += (double)xv[u] * (double)yv[u];
            }
    }
    store_partials_n<KMAX>(acc, K, partials);
}
__global__ __launch_bounds__(kBlock) void k_partials_fin(const double *partials, int B, double *result) {
    const double s = sum_partials(partials + (size_t)blockIdx.x * kPartLd, B);
    if (threadIdx.x == 0) result[blockIdx.x] = s;
}
template <typename T>
int launch_multidot(Ctx *c, int N, int K, const T *xs, size_t ldx, const T *y, double *result) {
    if (K <= 0) return SENK_OK;
    if (K > kRedMaxK) return SENK_ERR_UNSUPPORTED;
    const int grid = red_grid(N);
    double *P = red_slot(c, 6);
    if (K <= 2)
        k_multidot<T, 2><<<grid, kBlock, 0, c->stream>>>(N, K, xs, ldx, y, P);
    else if (K <= 8)
        k_multidot<T, 8><<<grid, kBlock, 0, c->stream>>>(N, K, xs, ldx, y, P);
    else
        k_multidot<T, kRedMaxK><<<grid, kBlock, 0, c->stream>>>(N, K, xs, ldx, y, P);
    SENK_LAUNCH_CHECK();
    k_partials_fin<<<K, kBlock, 0, c->stream>>>(P, grid, result);
    SENK_LAUNCH_CHECK();
    return SENK_OK;
}
template int launch_multidot<double>(Ctx *, int, int, const double *, size_t, const double *, double *);
template int launch_multidot<float>(Ctx *, int, int, const float *, size_t, const float *, double *);

__global__ __launch_bounds__(kBlock) void k_scaling(int N, const int *__restrict__ rptr, const int *__restrict__ cind,
                                                    double *__restrict__ val, const double *__restrict__ ld,
                                                    const double *__restrict__ rd) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= N) return;
    const double left = ld[i];
    for (int j = rptr[i]; j < rptr[i + 1]; j++) val[j] = left * val[j] * rd[cind[j]];
}
int launch_scaling_f64(Ctx *c, int N, const int *rptr, const int *cind, double *val, const double *ld,
                       const double *rd) {
    if (N <= 0) return SENK_OK;
    k_scaling<<<(N + kBlock - 1) / kBlock, kBlock, 0, c->stream>>>(N, rptr, cind, val, ld, rd);
    SENK_LAUNCH_CHECK();
    return SENK_OK;
}

}  // namespace senk_amd
