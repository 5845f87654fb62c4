// Triangular factors in a solve layout for the colored / level-scheduled SpTRSV.
//
// The rows one launch solves (the blocks of one schedule level, in schedule order) are
// cut into slices of 64 rows, one wave per slice, one lane per row; a slice therefore
// holds 64/BS whole blocks.  Every row's entries are stored in the order the
// sequential solve consumes them (senk_sptrsv.hpp: L ascending, U descending after the
// inverted diagonal), split in two groups:
//   external entries  column outside the row's block, solved by an earlier launch;
//                     value + column, k-th entry of every lane stored together;
//   in-block entries  column inside the block; value only, the column is the bit of a
//                     per-row mask, entries of step c of every lane stored together.
// Inside one group the k-th entries of the lanes that have one are packed with no
// gaps: a lane's position is the slice offset plus the entries of earlier k, plus the
// number of lower lanes that also have a k-th entry (wave ballot + mbcnt).  Loads are
// therefore contiguous runs, the stored volume is exactly nnz, and no padding exists.
// The arithmetic per row -- the operations and their order -- is that of the CSR
// kernels, so results are bit-identical to the sequential solve.
#include "dev_common.hpp"

#include <type_traits>

namespace senk_amd {

template <int BS>
using MaskT = typename std::conditional<(BS > 32), unsigned long long, unsigned>::type;

// number of set bits of b in the lanes below this one
__device__ __forceinline__ int lanes_below(unsigned long long b) {
    return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(b >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)b, 0u));
}

// Row of this lane: slice w of a level whose blocks are blk_list[blk0 .. blk0+nblk)
// (blk_list null: consecutive block numbers from blk0).
template <int BS>
struct LaneRow {
    int r, base, row;
    bool active;
    __device__ __forceinline__ LaneRow(int w, int blk0, int nblk, const int *blk_list) {
        const int lane = threadIdx.x & 63;
        const int lb = w * (64 / BS) + lane / BS;
        r = lane % BS;
        active = lb < nblk;
        base = active ? (blk_list ? blk_list[blk0 + lb] : blk0 + lb) * BS : 0;
        row = base + r;
    }
};

// Per row: number of external entries and in-block mask; per slice: entry counts.
// err is set when a row breaks the assumptions of the layout (strictly increasing
// columns; L strictly lower; U starting with its diagonal slot, then columns > row).
template <int BS, bool UPPER>
__global__ __launch_bounds__(kBlock) void k_tri_count(int slice0, int nsl, int blk0, int nblk,
                                                      const int *__restrict__ blk_list, const int *__restrict__ rptr,
                                                      const int *__restrict__ cind, int *__restrict__ next,
                                                      MaskT<BS> *__restrict__ mask, int *__restrict__ ecnt,
                                                      int *__restrict__ icnt, int *__restrict__ err) {
    const int w = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    if (w >= nsl) return;
    const LaneRow<BS> lr(w, blk0, nblk, blk_list);
    int ne = 0;
    MaskT<BS> mk = 0;
    bool bad = false;
    if (lr.active) {
        int s = rptr[lr.row];
        const int e = rptr[lr.row + 1];
        int prev = -1;
        if (UPPER) {
            if (e <= s || cind[s] != lr.row) bad = true;
            s++;
            prev = lr.row;
        }
        for (int j = s; j < e; j++) {
            const int col = cind[j];
            if (col <= prev) bad = true;
            prev = col;
            if (!UPPER) {
                if (col < lr.base) ne++;
                else if (col < lr.row) mk |= (MaskT<BS>)1 << (col - lr.base);
                else bad = true;
            } else {
                if (col >= lr.base + BS) ne++;
                else if (col > lr.row) mk |= (MaskT<BS>)1 << (col - lr.base);
                else bad = true;
            }
        }
    }
    const size_t slot = (size_t)(slice0 + w) * 64 + (threadIdx.x & 63);
    next[slot] = ne;
    mask[slot] = mk;
    const double se = wave_sum((double)ne), si = wave_sum((double)__popcll((unsigned long long)mk));
    if ((threadIdx.x & 63) == 0) {
        ecnt[slice0 + w] = (int)se;
        icnt[slice0 + w] = (int)si;
    }
    if (bad) *err = 1;
}

// Scatter one factor's values (and, once, the external columns) into the layout.
template <typename T, int BS, bool UPPER>
__global__ __launch_bounds__(kBlock) void k_tri_fill(int slice0, int nsl, int blk0, int nblk,
                                                     const int *__restrict__ blk_list, const int *__restrict__ rptr,
                                                     const int *__restrict__ cind, const T *__restrict__ val,
                                                     const int2 *__restrict__ soff, const int *__restrict__ next,
                                                     const MaskT<BS> *__restrict__ mask, int *__restrict__ scol,
                                                     T *__restrict__ ve, T *__restrict__ vi, T *__restrict__ dinv) {
    const int w = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    if (w >= nsl) return;
    const LaneRow<BS> lr(w, blk0, nblk, blk_list);
    const size_t slot = (size_t)(slice0 + w) * 64 + (threadIdx.x & 63);
    const int ne = next[slot];
    const MaskT<BS> mk = mask[slot];
    int s = 0, e = 0;
    if (lr.active) {
        s = rptr[lr.row];
        e = rptr[lr.row + 1];
    }
    const int2 so = soff[slice0 + w];
    int pos = so.x;
    for (int k = 0;; k++) {
        const unsigned long long b = __ballot(k < ne);
        if (!b) break;
        if (k < ne) {
            const int j = UPPER ? e - 1 - k : s + k;
            const int p = pos + lanes_below(b);
            if (scol) scol[p] = cind[j];
            ve[p] = val[j];
        }
        pos += __popcll(b);
    }
    int q = UPPER ? e - 1 - ne : s + ne;
    int ipos = so.y;
    for (int i = 0; i < BS; i++) {
        const int cc = UPPER ? BS - 1 - i : i;
        const bool on = (mk >> cc) & 1;
        const unsigned long long b = __ballot(on);
        if (on) {
            vi[ipos + lanes_below(b)] = val[q];
            q += UPPER ? -1 : 1;
        }
        ipos += __popcll(b);
    }
    if (UPPER) dinv[slot] = lr.active ? val[s] : T(0);
}

// One launch = one schedule level.  Phase 1 subtracts the external entries, CH gathers
// in flight per lane; phase 2 runs the BS in-block steps: at step c row c of every
// block is final and is broadcast inside the wave (U rows first scale by their
// inverted diagonal), and the rows holding column base+c subtract it.
template <typename T, int BS, bool UPPER>
__global__ __launch_bounds__(kBlock) void k_sptrsv_sell(int slice0, int nsl, int blk0, int nblk,
                                                        const int *__restrict__ blk_list,
                                                        const int2 *__restrict__ soff, const int *__restrict__ next,
                                                        const MaskT<BS> *__restrict__ mask,
                                                        const int *__restrict__ scol, const T *__restrict__ ve,
                                                        const T *__restrict__ vi, const T *__restrict__ dinv,
                                                        const T *x, T *y, const double *__restrict__ in_scale,
                                                        const int *__restrict__ stop_flag) {
    if (stop_flag && *stop_flag) return;
    const int w = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    if (w >= nsl) return;
    const LaneRow<BS> lr(w, blk0, nblk, blk_list);
    const size_t slot = (size_t)(slice0 + w) * 64 + (threadIdx.x & 63);
    const int ne = next[slot];
    const MaskT<BS> mk = mask[slot];
    const int2 so = soff[slice0 + w];
    const T dv = UPPER ? dinv[slot] : T(0);
    T t = 0;
    if (lr.active) {
        t = x[lr.row];
        if (!UPPER && in_scale) t = (T)((double)t * *in_scale);
    }
    constexpr int CH = 8;
    int pos = so.x;
    for (int k0 = 0; __ballot(k0 < ne); k0 += CH) {
        int idx[CH];
        bool on[CH];
#pragma unroll
        for (int q = 0; q < CH; q++) {
            on[q] = k0 + q < ne;
            const unsigned long long b = __ballot(on[q]);
            idx[q] = pos + lanes_below(b);
            pos += __popcll(b);
        }
        int cc[CH];
        T vv[CH], yv[CH];
#pragma unroll
        for (int q = 0; q < CH; q++)
            if (on[q]) {
                cc[q] = scol[idx[q]];
                vv[q] = ve[idx[q]];
            }
#pragma unroll
        for (int q = 0; q < CH; q++)
            if (on[q]) yv[q] = y[cc[q]];
#pragma unroll
        for (int q = 0; q < CH; q++)
            if (on[q]) t -= vv[q] * yv[q];
    }
    const int lbase = (threadIdx.x & 63) - lr.r;
    constexpr int IC = BS < 8 ? BS : 8;
    int ipos = so.y;
#pragma unroll
    for (int i0 = 0; i0 < BS; i0 += IC) {
        T iv[IC];
        bool on[IC];
#pragma unroll
        for (int q = 0; q < IC; q++) {
            const int c = UPPER ? BS - 1 - (i0 + q) : i0 + q;
            on[q] = (mk >> c) & 1;
            const unsigned long long b = __ballot(on[q]);
            iv[q] = on[q] ? vi[ipos + lanes_below(b)] : T(0);
            ipos += __popcll(b);
        }
#pragma unroll
        for (int q = 0; q < IC; q++) {
            const int c = UPPER ? BS - 1 - (i0 + q) : i0 + q;
            if (UPPER && lr.r == c) t = t * dv;
            if (BS > 1) {
                const T yc = __shfl(t, lbase + c, 64);
                if (on[q]) t -= iv[q] * yc;
            }
        }
    }
    if (lr.active) y[lr.row] = t;
}

// ----------------------------------------------------------------- host side
namespace {
inline int slices_of(int nblk, int bs) { return (int)(((long long)nblk * bs + 63) / 64); }
inline int waves_grid(int nsl) { return (nsl + kWavesPerBlock - 1) / kWavesPerBlock; }
bool layout_bsize_ok(int bs) { return bs >= 1 && bs <= 64 && (64 % bs) == 0; }

template <int BS>
int count_levels(Ctx *c, const Schedule &s, const TriLayout &L, const int *rptr, const int *cind, int *ecnt,
                 int *icnt, int *err) {
    for (int l = 0; l < s.nlev; l++) {
        const int b0 = s.lev_ptr[l], nb = s.lev_ptr[l + 1] - b0;
        const int sl0 = L.lev_slice[l], nsl = L.lev_slice[l + 1] - sl0;
        if (nsl <= 0) continue;
        if (L.upper)
            k_tri_count<BS, true><<<waves_grid(nsl), kBlock, 0, c->stream>>>(
                sl0, nsl, b0, nb, s.dblk, rptr, cind, L.next, (MaskT<BS> *)L.mask, ecnt, icnt, err);
        else
            k_tri_count<BS, false><<<waves_grid(nsl), kBlock, 0, c->stream>>>(
                sl0, nsl, b0, nb, s.dblk, rptr, cind, L.next, (MaskT<BS> *)L.mask, ecnt, icnt, err);
        SENK_LAUNCH_CHECK();
    }
    return SENK_OK;
}

template <typename T, int BS>
int fill_levels(Ctx *c, const Schedule &s, const TriLayout &L, const int *rptr, const int *cind, const T *val,
                int *scol, T *ve, T *vi, T *dinv) {
    for (int l = 0; l < s.nlev; l++) {
        const int b0 = s.lev_ptr[l], nb = s.lev_ptr[l + 1] - b0;
        const int sl0 = L.lev_slice[l], nsl = L.lev_slice[l + 1] - sl0;
        if (nsl <= 0) continue;
        if (L.upper)
            k_tri_fill<T, BS, true><<<waves_grid(nsl), kBlock, 0, c->stream>>>(
                sl0, nsl, b0, nb, s.dblk, rptr, cind, val, L.soff, L.next, (const MaskT<BS> *)L.mask, scol, ve, vi,
                dinv);
        else
            k_tri_fill<T, BS, false><<<waves_grid(nsl), kBlock, 0, c->stream>>>(
                sl0, nsl, b0, nb, s.dblk, rptr, cind, val, L.soff, L.next, (const MaskT<BS> *)L.mask, scol, ve, vi,
                dinv);
        SENK_LAUNCH_CHECK();
    }
    return SENK_OK;
}

template <typename T, int BS>
int solve_levels(Ctx *c, const Schedule &s, const TriLayout &L, const T *ve, const T *vi, const T *dinv, const T *x,
                 T *y, const double *in_scale, const int *stop) {
    for (int step = 0; step < s.nlev; step++) {
        const int l = L.upper ? s.nlev - 1 - step : step;
        const int b0 = s.lev_ptr[l], nb = s.lev_ptr[l + 1] - b0;
        const int sl0 = L.lev_slice[l], nsl = L.lev_slice[l + 1] - sl0;
        if (nsl <= 0) continue;
        if (L.upper)
            k_sptrsv_sell<T, BS, true><<<waves_grid(nsl), kBlock, 0, c->stream>>>(
                sl0, nsl, b0, nb, s.dblk, L.soff, L.next, (const MaskT<BS> *)L.mask, L.scol, ve, vi, dinv, x, y,
                nullptr, stop);
        else
            k_sptrsv_sell<T, BS, false><<<waves_grid(nsl), kBlock, 0, c->stream>>>(
                sl0, nsl, b0, nb, s.dblk, L.soff, L.next, (const MaskT<BS> *)L.mask, L.scol, ve, vi, dinv, x, y,
                in_scale, stop);
        SENK_LAUNCH_CHECK();
    }
    return SENK_OK;
}

#define SENK_BS_DISPATCH(BSV, CALL)        \
    switch (BSV) {                         \
        case 1: { constexpr int B = 1; return CALL; }   \
        case 2: { constexpr int B = 2; return CALL; }   \
        case 4: { constexpr int B = 4; return CALL; }   \
        case 8: { constexpr int B = 8; return CALL; }   \
        case 16: { constexpr int B = 16; return CALL; } \
        case 32: { constexpr int B = 32; return CALL; } \
        case 64: { constexpr int B = 64; return CALL; } \
        default: return SENK_ERR_UNSUPPORTED;           \
    }
}  // namespace

void tri_layout_free(TriLayout *L) {
    for (void *p : {(void *)L->soff, (void *)L->next, L->mask, (void *)L->scol, (void *)L->ve64, (void *)L->vi64,
                    (void *)L->d64, (void *)L->ve32, (void *)L->vi32, (void *)L->d32})
        if (p) (void)hipFree(p);
    *L = TriLayout{};
}

int tri_layout_build(Ctx *c, const Schedule &s, bool upper, int N, const int *rptr, const int *cind,
                     TriLayout *L) {
    const int BS = s.bsize;
    if (!layout_bsize_ok(BS) || (long long)s.lev_ptr.size() < (long long)s.nlev + 1) return SENK_ERR_UNSUPPORTED;
    tri_layout_free(L);
    L->bsize = BS;
    L->upper = upper;
    L->lev_slice.assign((size_t)s.nlev + 1, 0);
    for (int l = 0; l < s.nlev; l++)
        L->lev_slice[l + 1] = L->lev_slice[l] + slices_of(s.lev_ptr[l + 1] - s.lev_ptr[l], BS);
    L->nslices = L->lev_slice[s.nlev];
    if ((long long)s.lev_ptr[s.nlev] * BS != N) return SENK_ERR_SCHEDULE;
    const size_t nslots = (size_t)L->nslices * 64;
    const size_t mbytes = BS > 32 ? sizeof(unsigned long long) : sizeof(unsigned);
    int *ecnt = nullptr, *icnt = nullptr, *err = nullptr;
    int st = SENK_OK;
    std::vector<int> he(L->nslices), hi(L->nslices);
    int herr = 0;
    auto fail = [&](int code) {
        (void)hipFree(ecnt);
        (void)hipFree(icnt);
        (void)hipFree(err);
        if (code) tri_layout_free(L);
        return code;
    };
    if (L->nslices == 0) return SENK_OK;
    if (hipMalloc(&L->next, nslots * sizeof(int)) != hipSuccess || hipMalloc(&L->mask, nslots * mbytes) != hipSuccess ||
        hipMalloc(&ecnt, L->nslices * sizeof(int)) != hipSuccess ||
        hipMalloc(&icnt, L->nslices * sizeof(int)) != hipSuccess || hipMalloc(&err, sizeof(int)) != hipSuccess)
        return fail(SENK_ERR_ALLOC);
    if (hipMemsetAsync(err, 0, sizeof(int), c->stream) != hipSuccess) return fail(SENK_ERR_HIP);
    auto count = [&]() -> int { SENK_BS_DISPATCH(BS, count_levels<B>(c, s, *L, rptr, cind, ecnt, icnt, err)) };
    if ((st = count())) return fail(st);
    if (hipMemcpyAsync(he.data(), ecnt, he.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipMemcpyAsync(hi.data(), icnt, hi.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return fail(SENK_ERR_HIP);
    if (herr) return fail(SENK_ERR_UNSUPPORTED);
    std::vector<int2> so(L->nslices);
    long long ne = 0, ni = 0;
    for (int i = 0; i < L->nslices; i++) {
        so[i] = make_int2((int)ne, (int)ni);
        ne += he[i];
        ni += hi[i];
    }
    if (ne > INT32_MAX || ni > INT32_MAX) return fail(SENK_ERR_UNSUPPORTED);
    L->n_ext = ne;
    L->n_int = ni;
    if (hipMalloc(&L->soff, so.size() * sizeof(int2)) != hipSuccess ||
        hipMalloc(&L->scol, (size_t)(ne ? ne : 1) * sizeof(int)) != hipSuccess)
        return fail(SENK_ERR_ALLOC);
    if (hipMemcpyAsync(L->soff, so.data(), so.size() * sizeof(int2), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return fail(SENK_ERR_HIP);
    L->cols_filled = false;
    return fail(SENK_OK);
}

template <typename T>
int tri_layout_fill(Ctx *c, const Schedule &s, TriLayout *L, const int *rptr, const int *cind, const T *val) {
    if (L->nslices == 0) return SENK_OK;
    T **ve = std::is_same<T, double>::value ? (T **)&L->ve64 : (T **)&L->ve32;
    T **vi = std::is_same<T, double>::value ? (T **)&L->vi64 : (T **)&L->vi32;
    T **dv = std::is_same<T, double>::value ? (T **)&L->d64 : (T **)&L->d32;
    if (!*ve) SENK_HIP(hipMalloc(ve, (size_t)(L->n_ext ? L->n_ext : 1) * sizeof(T)));
    if (!*vi) SENK_HIP(hipMalloc(vi, (size_t)(L->n_int ? L->n_int : 1) * sizeof(T)));
    if (L->upper && !*dv) SENK_HIP(hipMalloc(dv, (size_t)L->nslices * 64 * sizeof(T)));
    int *scol = L->cols_filled ? nullptr : L->scol;
    auto run = [&]() -> int { SENK_BS_DISPATCH(L->bsize, (fill_levels<T, B>(c, s, *L, rptr, cind, val, scol, *ve, *vi, *dv))) };
    SENK_TRY(run());
    L->cols_filled = true;
    return SENK_OK;
}
template int tri_layout_fill<double>(Ctx *, const Schedule &, TriLayout *, const int *, const int *, const double *);
template int tri_layout_fill<float>(Ctx *, const Schedule &, TriLayout *, const int *, const int *, const float *);

template <typename T>
int launch_sptrsv_layout(Ctx *c, const Schedule &s, const TriLayout &L, const T *x, T *y, const double *in_scale,
                         const int *stop) {
    const T *ve = std::is_same<T, double>::value ? (const T *)L.ve64 : (const T *)L.ve32;
    const T *vi = std::is_same<T, double>::value ? (const T *)L.vi64 : (const T *)L.vi32;
    const T *dv = std::is_same<T, double>::value ? (const T *)L.d64 : (const T *)L.d32;
    if (L.nslices == 0) return SENK_OK;
    if (!ve || !vi || (L.upper && !dv) || L.bsize != s.bsize) return SENK_ERR_ARG;
    SENK_BS_DISPATCH(L.bsize, (solve_levels<T, B>(c, s, L, ve, vi, dv, x, y, in_scale, stop)))
}
template int launch_sptrsv_layout<double>(Ctx *, const Schedule &, const TriLayout &, const double *, double *,
                                          const double *, const int *);
template int launch_sptrsv_layout<float>(Ctx *, const Schedule &, const TriLayout &, const float *, float *,
                                         const double *, const int *);

}  // namespace senk_amd
