// Triangular factors in a solve layout for the colored / level-scheduled SpTRSV.
//
// The rows of one schedule level (its blocks, in schedule order) are cut into slices
// of 64 rows, one wave per slice, one lane per row; a slice therefore holds 64/BS whole
// blocks.  Slices are numbered in solve order (L: levels ascending, U: descending).
// Every row's entries are stored in the order the sequential solve consumes them
// (senk_sparse.hpp:133-167: L ascending, U descending after the inverted diagonal), split in
// two groups:
//   external entries  column outside the row's block, solved in an earlier level;
//                     value + column, k-th entry of every lane stored together;
//   in-block entries  column inside the block; value only, the column is the bit of a
//                     per-row mask, entries of step c of every lane stored together.
// Inside one group the k-th entries of the lanes that have one are packed with no
// gaps: a lane's position is the slice offset plus the entries of earlier k, plus the
// number of lower lanes that also have a k-th entry (wave ballot + mbcnt).  Loads are
// therefore contiguous runs, the stored volume is exactly nnz, and no padding exists.
// Per slice one 16-byte descriptor (entry offsets, first block, block count) and per
// row one word (external count and in-block mask) are all a wave reads before it can
// address its entries.  The arithmetic per row -- the operations and their order -- is
// that of the CSR kernels, so results are bit-identical to the sequential solve.
#include "dev_common.hpp"

#include <hip/hip_fp16.h>

#include <algorithm>
#include <type_traits>

namespace senk_amd {

template <int BS>
using MaskT = typename std::conditional<(BS > 32), unsigned long long, unsigned>::type;

// Per-row word: BS <= 16 external count << 16 | mask; BS == 32 count << 32 | mask;
// BS == 64 the mask alone (the count in a separate array).
template <int BS>
struct RowWord {
    static constexpr bool kPacked = BS <= 32;
    using W = typename std::conditional<(BS <= 16), unsigned, unsigned long long>::type;
    static constexpr int kShift = BS <= 16 ? 16 : 32;
    static constexpr long long kMaxExt = BS <= 16 ? 0xffff : 0x7fffffff;
    __device__ __forceinline__ static W pack(int ne, MaskT<BS> mk) {
        if constexpr (kPacked) return ((W)ne << kShift) | (W)mk;
        else return (W)mk;
    }
    __device__ __forceinline__ static void unpack(W w, const int *next, size_t slot, int &ne, MaskT<BS> &mk) {
        if constexpr (kPacked) {
            ne = (int)(w >> kShift);
            mk = (MaskT<BS>)(w & (((W)1 << kShift) - 1));
        } else {
            ne = next[slot];
            mk = (MaskT<BS>)w;
        }
    }
};

// Stored factor values: FP64 / FP32 as computed with, or FP16 holding value * 2^-e with
// one exponent e per row (|stored| < 1; FP32 arithmetic on 2^e * stored, which is exact).
template <typename V>
constexpr bool kHalfV = std::is_same<V, __half>::value;
template <typename T, typename V>
__device__ __forceinline__ T decode(V v, int e) {
    if constexpr (kHalfV<V>) return ldexpf(__half2float(v), e);
    else return (T)v;
}
// inverted diagonal stored with FP16 factors: FP32
template <typename V>
using DiagT = typename std::conditional<kHalfV<V>, float, V>::type;

// number of set bits of b in the lanes below this one
__device__ __forceinline__ int lanes_below(unsigned long long b) {
    return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(b >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)b, 0u));
}

// Row c of every block of BS consecutive lanes, broadcast to the block's lanes (the
// source row of an in-block step), without the LDS crossbar of a general shuffle: DPP
// inside rows of 16 lanes (quad_perm for blocks of 2 and 4, row_newbcast for 8 and 16)
// and v_readlane for 32 and 64.  c is a compile-time constant once the in-block loop is
// unrolled, so each switch folds to one instruction.  Every lane of the wave must be
// active (the in-block loop runs in uniform control flow).
template <int CTRL>
__device__ __forceinline__ int dpp_mov(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, false);
}
__device__ __forceinline__ int row_bcast16(int v, int n) {
    switch (n) {
        case 0: return dpp_mov<0x150>(v);
        case 1: return dpp_mov<0x151>(v);
        case 2: return dpp_mov<0x152>(v);
        case 3: return dpp_mov<0x153>(v);
        case 4: return dpp_mov<0x154>(v);
        case 5: return dpp_mov<0x155>(v);
        case 6: return dpp_mov<0x156>(v);
        case 7: return dpp_mov<0x157>(v);
        case 8: return dpp_mov<0x158>(v);
        case 9: return dpp_mov<0x159>(v);
        case 10: return dpp_mov<0x15a>(v);
        case 11: return dpp_mov<0x15b>(v);
        case 12: return dpp_mov<0x15c>(v);
        case 13: return dpp_mov<0x15d>(v);
        case 14: return dpp_mov<0x15e>(v);
        default: return dpp_mov<0x15f>(v);
    }
}
template <int BS>
__device__ __forceinline__ int block_bcast32(int v, int c) {
    const int lane = threadIdx.x & 63;
    if constexpr (BS == 2) {  // quad_perm [c, c, 2 + c, 2 + c]
        return c == 0 ? dpp_mov<0xa0>(v) : dpp_mov<0xf5>(v);
    } else if constexpr (BS == 4) {  // quad_perm [c, c, c, c]
        switch (c) {
            case 0: return dpp_mov<0x00>(v);
            case 1: return dpp_mov<0x55>(v);
            case 2: return dpp_mov<0xaa>(v);
            default: return dpp_mov<0xff>(v);
        }
    } else if constexpr (BS == 8) {
        const int a = row_bcast16(v, c), b = row_bcast16(v, c + 8);
        return (lane & 8) ? b : a;
    } else if constexpr (BS == 16) {
        return row_bcast16(v, c);
    } else if constexpr (BS == 32) {
        const int a = __builtin_amdgcn_readlane(v, c), b = __builtin_amdgcn_readlane(v, 32 + c);
        return lane < 32 ? a : b;
    } else {
        return __builtin_amdgcn_readlane(v, c);
    }
}
template <int BS, typename T>
__device__ __forceinline__ T block_bcast(T v, int c) {
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, block_bcast32<BS>(__builtin_bit_cast(int, v), c));
    } else {
        const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
        const unsigned lo = (unsigned)block_bcast32<BS>((int)(unsigned)u, c);
        const unsigned hi = (unsigned)block_bcast32<BS>((int)(unsigned)(u >> 32), c);
        return __builtin_bit_cast(T, ((unsigned long long)hi << 32) | lo);
    }
}

// Row of this lane in a slice whose descriptor is si = {external offset, in-block
// offset, first block entry, blocks}: blocks blk_list[si.z ..] (blk_list null:
// consecutive block numbers from si.z).
template <int BS>
struct LaneRow {
    int r, base, row;
    bool active;
    __device__ __forceinline__ LaneRow(int4 si, const int *blk_list) {
        const int lane = threadIdx.x & 63;
        const int lb = lane / BS;
        r = lane % BS;
        active = lb < si.w;
        base = active ? (blk_list ? blk_list[si.z + lb] : si.z + lb) * BS : 0;
        row = base + r;
    }
};

// Per row: the row word; per slice: entry counts (ecnt[w] external entries of slice w,
// ecnt[nslices + w] its rows' largest external count).  err = 1 when a row breaks the
// assumptions of the layout (strictly increasing columns; L strictly lower; U starting
// with its diagonal slot, then columns > row), 2 when a count does not fit the word.
template <int BS, bool UPPER>
__global__ __launch_bounds__(kBlock) void k_tri_count(int nslices, const int4 *__restrict__ sinfo,
                                                      const int *__restrict__ blk_list, const int *__restrict__ rptr,
                                                      const int *__restrict__ cind,
                                                      typename RowWord<BS>::W *__restrict__ rw, int *__restrict__ next,
                                                      int *__restrict__ ecnt, int *__restrict__ icnt,
                                                      int *__restrict__ err) {
    const int w = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    if (w >= nslices) return;
    const LaneRow<BS> lr(sinfo[w], blk_list);
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
    const bool wide = ne > RowWord<BS>::kMaxExt;
    const size_t slot = (size_t)w * 64 + (threadIdx.x & 63);
    rw[slot] = RowWord<BS>::pack(wide ? 0 : ne, mk);
    if (!RowWord<BS>::kPacked) next[slot] = ne;
    const double se = wave_sum((double)ne), si = wave_sum((double)__popcll((unsigned long long)mk));
    int mx = ne;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mx = max(mx, __shfl_xor(mx, off, 64));
    if ((threadIdx.x & 63) == 0) {
        ecnt[w] = (int)se;
        ecnt[nslices + w] = mx;
        icnt[w] = (int)si;
    }
    if (bad) *err = 1;
    else if (wide) *err = 2;
}

// Scatter one factor's values (and, once, the external columns) into the layout.  FP16
// storage: the row exponent e is the binary exponent of the row's largest magnitude
// (the U diagonal slot excluded), so every stored value lies in (-1, 1).
template <typename S, typename V, int BS, bool UPPER>
__global__ __launch_bounds__(kBlock) void k_tri_fill(int nslices, const int4 *__restrict__ sinfo,
                                                     const int *__restrict__ blk_list, const int *__restrict__ rptr,
                                                     const int *__restrict__ cind, const S *__restrict__ val,
                                                     const typename RowWord<BS>::W *__restrict__ rw,
                                                     const int *__restrict__ next, int *__restrict__ scol,
                                                     V *__restrict__ ve, V *__restrict__ vi,
                                                     DiagT<V> *__restrict__ dinv, signed char *__restrict__ sx) {
    const int w = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    if (w >= nslices) return;
    const int4 si = sinfo[w];
    const LaneRow<BS> lr(si, blk_list);
    const size_t slot = (size_t)w * 64 + (threadIdx.x & 63);
    int ne;
    MaskT<BS> mk;
    RowWord<BS>::unpack(rw[slot], next, slot, ne, mk);
    int s = 0, e = 0;
    if (lr.active) {
        s = rptr[lr.row];
        e = rptr[lr.row + 1];
    }
    int ex = 0;
    if constexpr (kHalfV<V>) {
        double m = 0;
        for (int j = UPPER ? s + 1 : s; j < e; j++) m = fmax(m, fabs((double)val[j]));
        if (m > 0) (void)frexp(m, &ex);
        sx[slot] = (signed char)ex;
    }
    auto enc = [&](S v) -> V {
        if constexpr (kHalfV<V>) return __float2half((float)ldexp((double)v, -ex));
        else return (V)v;
    };
    int pos = si.x;
    for (int k = 0;; k++) {
        const unsigned long long b = __ballot(k < ne);
        if (!b) break;
        if (k < ne) {
            const int j = UPPER ? e - 1 - k : s + k;
            const int p = pos + lanes_below(b);
            if (scol) scol[p] = cind[j];
            ve[p] = enc(val[j]);
        }
        pos += __popcll(b);
    }
    int q = UPPER ? e - 1 - ne : s + ne;
    int ipos = si.y;
    for (int i = 0; i < BS; i++) {
        const int cc = UPPER ? BS - 1 - i : i;
        const bool on = (mk >> cc) & 1;
        const unsigned long long b = __ballot(on);
        if (on) {
            vi[ipos + lanes_below(b)] = enc(val[q]);
            q += UPPER ? -1 : 1;
        }
        ipos += __popcll(b);
    }
    if (UPPER) dinv[slot] = lr.active ? (DiagT<V>)val[s] : DiagT<V>(0);
}

// ------------------------------------------------------------ single-launch sweep
// In the sweep every row of the output starts as a sentinel: a signalling NaN, which
// no arithmetic produces (results are canonicalised before they are stored, so an
// input holding the pattern cannot leak through either).  A gather of an earlier row
// that still reads the sentinel polls again; a row's value is stored once, so the
// first non-sentinel value read is the final one.  Stores of the output and the polls
// go to the device-coherent level (every XCD has its own L2, which a row written on
// another XCD does not update).
template <typename T>
struct Bits {
    using U = unsigned;
    static constexpr U kSentinel = 0x7fa5a5a5u;
};
template <>
struct Bits<double> {
    using U = unsigned long long;
    static constexpr U kSentinel = 0x7ff4a5a5a5a5a5a5ull;
};
template <typename T>
__device__ __forceinline__ typename Bits<T>::U load_agent(const T *p) {
    using U = typename Bits<T>::U;
    return __hip_atomic_load((const U *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// First read of a gathered row: through the XCD's L2 (workgroup scope).  At the start of
// a launch no cache holds an older value of any row (the sentinels were written by an
// earlier launch), and within the launch a row changes once, from the sentinel to its
// final value; so this read gives the final value or the sentinel, and only a sentinel
// is polled again at device scope.
template <typename T>
__device__ __forceinline__ typename Bits<T>::U load_local(const T *p) {
    using U = typename Bits<T>::U;
    return __hip_atomic_load((const U *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <typename T>
__device__ __forceinline__ void store_agent(T *p, T v) {
    using U = typename Bits<T>::U;
    __hip_atomic_store((U *)p, __builtin_bit_cast(U, v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float canonical(float v) { return __builtin_canonicalizef(v); }
__device__ __forceinline__ double canonical(double v) { return __builtin_canonicalize(v); }
// p + i; with O32 addressed as p's (uniform) base plus a 32-bit byte offset, one address
// register per access instead of a 64-bit pair (the launch checks that every array of the
// layout and every vector fits: i * sizeof(P) < 2^32)
template <bool O32, typename P>
__device__ __forceinline__ P *ixp(P *p, size_t i) {
    if constexpr (O32) {
        using Ch = typename std::conditional<std::is_const<P>::value, const char, char>::type;
        return reinterpret_cast<P *>(reinterpret_cast<Ch *>(p) + (size_t)((unsigned)i * (unsigned)sizeof(P)));
    } else {
        return p + i;
    }
}

// Sweep words (unsigned, device): the time-out flag and the ticket counter of mode 2.
constexpr int kSweepErr = 0, kSweepTicket = 1, kSweepWords = 32;
// workgroup size (waves = slices per ticket) of the per-XCD ticket streams: 8-wave
// workgroups measured best (FP32 application at 128^3, connect blocks of 8, with one
// ticket counter per XCD: 71.5 us, 77.4 with 16 waves, 73.0 with 4)
constexpr int kXTicketWaves = 8;
// polls of one pending gather before the wave gives up and raises the time-out flag;
// a poll takes about a microsecond, so this bounds a wait at a few seconds
constexpr int kSweepSpinLimit = 1 << 22;

// The rows of one slice.  The wave's own data is requested first (descriptor, row
// words, right-hand side, first chunk of external entries, first in-block values);
// phase 1 then subtracts the external entries, CH gathers in flight per lane; phase 2
// runs the BS in-block steps: at step c row c of every block is final and is broadcast
// inside the wave (U rows first scale by their inverted diagonal), and the rows holding
// column base+c subtract it.  SWEEP: gathers wait on the sentinel, the row is stored
// device-coherent, and reset[row] (when given) receives the sentinel once x[row] has
// been read.
// What a wave reads about its slice before it can address the slice's entries: the
// descriptor and, per lane, the row word (external count, in-block mask), the inverted
// diagonal (U) and the FP16 row exponent.
template <typename V, int BS>
struct SliceHead {
    int4 si;
    int ne;
    MaskT<BS> mk;
    DiagT<V> dv;
    int ex;
};
template <typename V, int BS, bool UPPER, bool O32 = false>
__device__ __forceinline__ SliceHead<V, BS> slice_head(int slice, const int4 *__restrict__ sinfo,
                                                       const typename RowWord<BS>::W *__restrict__ rw,
                                                       const int *__restrict__ next,
                                                       const DiagT<V> *__restrict__ dinv,
                                                       const signed char *__restrict__ sx) {
    SliceHead<V, BS> h;
    h.si = *ixp<O32>(sinfo, slice);
    const size_t slot = (size_t)slice * 64 + (threadIdx.x & 63);
    RowWord<BS>::unpack(*ixp<O32>(rw, slot), next, slot, h.ne, h.mk);
    h.dv = UPPER ? *ixp<O32>(dinv, slot) : DiagT<V>(0);
    h.ex = kHalfV<V> ? (int)*ixp<O32>(sx, slot) : 0;
    return h;
}

template <typename T, typename V, int BS, bool UPPER, bool SWEEP, int CH = 8, bool PIPE = false, bool O32 = false>
__device__ __forceinline__ void solve_slice_body(const SliceHead<V, BS> &h, const int *__restrict__ blk_list,
                                                 const int *__restrict__ scol, const V *__restrict__ ve,
                                                 const V *__restrict__ vi, const T *x, T *y,
                                                 const double *__restrict__ in_scale, T *reset, unsigned *err,
                                                 T *copy, const T *ep_a, T *ep_b,
                                                 const SweepPre<T> &pre = SweepPre<T>{}) {
    using U = typename Bits<T>::U;
    const int4 si = h.si;
    const LaneRow<BS> lr(si, blk_list);
    const int ne = h.ne;
    const MaskT<BS> mk = h.mk;
    const T dv = UPPER ? (T)h.dv : T(0);
    const int ex = h.ex;
    T t = 0, pa = 0, pb = 0;
    if (lr.active) {
        t = *ixp<O32>(x, lr.row);
        if constexpr (UPPER && SWEEP) {
            if (pre.u) {  // the right-hand side's update (x == pre.u), all loads first
                const T a = pre.ab[0], be = pre.ab[1];
                const T r0 = *ixp<O32>(pre.r, lr.row), qv = *ixp<O32>(pre.q, lr.row),
                        yv = *ixp<O32>(pre.yd, lr.row);
                const T rv = (-a) * qv + r0;
                *ixp<O32>(pre.yd, lr.row) = a * t + yv;
                *ixp<O32>(pre.r, lr.row) = rv;
                t = rv + be * t;
                *ixp<O32>(pre.u, lr.row) = t;
            }
        }
        if (!UPPER && in_scale) t = (T)((double)t * *in_scale);
        // requested with the right-hand side, used at the end; read before the reset, as
        // ep_b may be the U sweep's own vector, which this L sweep returns to the sentinel
        if (!UPPER && SWEEP && ep_a) {
            pa = *ixp<O32>(ep_a, lr.row);
            pb = *ixp<O32>(ep_b, lr.row);
        }
        if (SWEEP && reset) *ixp<O32>(reset, lr.row) = __builtin_bit_cast(T, Bits<T>::kSentinel);
    }
    const T rhs = UPPER ? t : T(0);  // kept for ep_b (U only)
    // CH external entries per lane per round of gathers
    constexpr int IC = BS < 8 ? BS : 8;
    int pos = si.x, k0 = 0;
    int cc[CH];
    bool on[CH];
    T vv[CH];
    bool more = __ballot(0 < ne) != 0;
    // entries kk .. kk+CH-1 of every lane (the packed positions advance in entry order)
    auto load_chunk_at = [&](int kk, int (&c_)[CH], bool (&o_)[CH], T (&v_)[CH]) {
        int idx[CH];
#pragma unroll
        for (int q = 0; q < CH; q++) {
            o_[q] = kk + q < ne;
            const unsigned long long b = __ballot(o_[q]);
            idx[q] = pos + lanes_below(b);
            pos += __popcll(b);
        }
#pragma unroll
        for (int q = 0; q < CH; q++)
            if (o_[q]) {
                c_[q] = *ixp<O32>(scol, idx[q]);
                v_[q] = decode<T>(*ixp<O32>(ve, idx[q]), ex);
            }
    };
    auto load_chunk = [&]() { load_chunk_at(k0, cc, on, vv); };
    if (more) load_chunk();
    T iv0[IC];
    bool on0[IC];
    int ipos = si.y;
#pragma unroll
    for (int q = 0; q < IC; q++) {
        const int c = UPPER ? BS - 1 - q : q;
        on0[q] = (mk >> c) & 1;
        const unsigned long long b = __ballot(on0[q]);
        iv0[q] = on0[q] ? decode<T>(*ixp<O32>(vi, ipos + lanes_below(b)), ex) : T(0);
        ipos += __popcll(b);
    }
    while (more) {
        T yv[CH];
        // PIPE (rows with more than CH external entries): the next round's entries are
        // requested while this round's gathers are in flight, so a round costs one
        // dependent memory round trip instead of two
        int ccn[CH];
        bool onn[CH];
        T vvn[CH];
        bool moren = false;
        if (SWEEP) {
            U yb[CH];
            bool pend = false;
#pragma unroll
            for (int q = 0; q < CH; q++)
                if (on[q]) yb[q] = load_local(ixp<O32>(y, cc[q]));
            if constexpr (PIPE) {
                moren = __ballot(k0 + CH < ne) != 0;
                if (moren) load_chunk_at(k0 + CH, ccn, onn, vvn);
            }
#pragma unroll
            for (int q = 0; q < CH; q++) pend |= on[q] && yb[q] == Bits<T>::kSentinel;
            for (int spin = 0; __ballot(pend); spin++) {
                if (spin >= kSweepSpinLimit) {
                    if (pend) __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
                pend = false;
#pragma unroll
                for (int q = 0; q < CH; q++)
                    if (on[q] && yb[q] == Bits<T>::kSentinel) {
                        yb[q] = load_agent(ixp<O32>(y, cc[q]));
                        pend |= yb[q] == Bits<T>::kSentinel;
                    }
            }
#pragma unroll
            for (int q = 0; q < CH; q++) yv[q] = __builtin_bit_cast(T, yb[q]);
        } else {
#pragma unroll
            for (int q = 0; q < CH; q++)
                if (on[q]) yv[q] = y[cc[q]];
        }
#pragma unroll
        for (int q = 0; q < CH; q++)
            if (on[q]) t -= vv[q] * yv[q];
        k0 += CH;
        if constexpr (PIPE && SWEEP) {
            more = moren;
            if (more) {
#pragma unroll
                for (int q = 0; q < CH; q++) {
                    cc[q] = ccn[q];
                    on[q] = onn[q];
                    vv[q] = vvn[q];
                }
            }
        } else {
            more = __ballot(k0 < ne) != 0;
            if (more) load_chunk();
        }
    }
#pragma unroll
    for (int i0 = 0; i0 < BS; i0 += IC) {
        T iv[IC];
        bool onq[IC];
#pragma unroll
        for (int q = 0; q < IC; q++) {
            if (i0 == 0) {
                onq[q] = on0[q];
                iv[q] = iv0[q];
            } else {
                const int c = UPPER ? BS - 1 - (i0 + q) : i0 + q;
                onq[q] = (mk >> c) & 1;
                const unsigned long long b = __ballot(onq[q]);
                iv[q] = onq[q] ? decode<T>(*ixp<O32>(vi, ipos + lanes_below(b)), ex) : T(0);
                ipos += __popcll(b);
            }
        }
#pragma unroll
        for (int q = 0; q < IC; q++) {
            const int c = UPPER ? BS - 1 - (i0 + q) : i0 + q;
            if (UPPER && lr.r == c) t = t * dv;
            if (BS > 1) {
                const T yc = block_bcast<BS>(t, c);
                if (onq[q]) t -= iv[q] * yc;
            }
        }
    }
    if (lr.active) {
        if (SWEEP) {
            store_agent(ixp<O32>(y, lr.row), canonical(t));
            // optional epilogue vectors (single-launch modes): a U sweep also writes
            // ep_b = rhs - ep_a * t (the vector pass that follows a U solve in the split
            // forms); an L sweep writes ep_a * ep_b + t instead of t to copy (the split
            // CG's product D t + L^-1 w, ep_a = D, ep_b = t)
            if constexpr (UPPER) {
                if (copy) *ixp<O32>(copy, lr.row) = t;
                if (ep_b) *ixp<O32>(ep_b, lr.row) = rhs - *ixp<O32>(ep_a, lr.row) * t;
            } else {
                if (copy) *ixp<O32>(copy, lr.row) = ep_a ? pa * pb + t : t;
            }
        } else {
            y[lr.row] = t;
        }
    }
}

template <typename T, typename V, int BS, bool UPPER, bool SWEEP>
__device__ __forceinline__ void solve_slice(int slice, const int4 *__restrict__ sinfo, const int *__restrict__ blk_list,
                                            const typename RowWord<BS>::W *__restrict__ rw,
                                            const int *__restrict__ next, const int *__restrict__ scol,
                                            const V *__restrict__ ve, const V *__restrict__ vi,
                                            const DiagT<V> *__restrict__ dinv, const signed char *__restrict__ sx,
                                            const T *x, T *y, const double *__restrict__ in_scale, T *reset,
                                            unsigned *err, T *copy = nullptr, const T *ep_a = nullptr,
                                            T *ep_b = nullptr, const SweepPre<T> &pre = SweepPre<T>{}) {
    solve_slice_body<T, V, BS, UPPER, SWEEP>(slice_head<V, BS, UPPER>(slice, sinfo, rw, next, dinv, sx), blk_list,
                                             scol, ve, vi, x, y, in_scale, reset, err, copy, ep_a, ep_b, pre);
}

// One launch = one schedule level (slices slice0 .. slice0 + nsl).
template <typename T, typename V, int BS, bool UPPER>
__global__ __launch_bounds__(kBlock) void k_sptrsv_sell(int slice0, int nsl, int tord, const int4 *__restrict__ sinfo,
                                                        const int *__restrict__ blk_list,
                                                        const typename RowWord<BS>::W *__restrict__ rw,
                                                        const int *__restrict__ next, const int *__restrict__ scol,
                                                        const V *__restrict__ ve, const V *__restrict__ vi,
                                                        const DiagT<V> *__restrict__ dinv,
                                                        const signed char *__restrict__ sx, const T *x, T *y,
                                                        const double *__restrict__ in_scale,
                                                        const int *__restrict__ stop_flag) {
    if (stop_flag && *stop_flag) return;
    // each XCD takes one band of the level: within one color, neighbouring slices cover
    // neighbouring parts of the domain, so their gathers share that L2.  In slice order
    // the band is a contiguous part of the level; in ticket order (tord) XCD x holds
    // slices 64k + 8x .. 64k + 8x + 7 of every group of 64, two workgroups per 8 slices.
    int w;
    if (tord) {
        const int b = (int)blockIdx.x, i = b / kXcds;
        w = (i >> 1) * (kXcds * kXTicketWaves) + (b % kXcds) * kXTicketWaves + (i & 1) * kWavesPerBlock +
            (int)(threadIdx.x >> 6);
    } else {
        w = xcd_block() * kWavesPerBlock + (int)(threadIdx.x >> 6);
    }
    if (w >= nsl) return;
    solve_slice<T, V, BS, UPPER, false>(slice0 + w, sinfo, blk_list, rw, next, scol, ve, vi, dinv, sx, x, y,
                                        in_scale, nullptr, nullptr);
}

// The whole triangle in one cooperative launch: every wave of the grid is resident at
// once, and wave g solves slices g, g + P, g + 2P, ... (P waves in the grid).  A gather
// only waits on rows of earlier levels, i.e. of smaller slice numbers; the smallest
// unfinished slice therefore waits on nothing unfinished and its wave, being resident
// and done with its own smaller slices, works on it: the sweep always drains.  y must
// hold the sentinel in every row on entry.
template <typename T, typename V, int BS, bool UPPER>
__global__ __launch_bounds__(kBlock) void k_sptrsv_sweep(int nslices, const int4 *__restrict__ sinfo,
                                                         const int *__restrict__ blk_list,
                                                         const typename RowWord<BS>::W *__restrict__ rw,
                                                         const int *__restrict__ next, const int *__restrict__ scol,
                                                         const V *__restrict__ ve, const V *__restrict__ vi,
                                                         const DiagT<V> *__restrict__ dinv,
                                                         const signed char *__restrict__ sx, const T *x, T *y,
                                                         const double *__restrict__ in_scale, T *reset,
                                                         unsigned *err, T *copy, const T *ep_a, T *ep_b,
                                                         const int *__restrict__ stop_flag, SweepPre<T> pre) {
    if (stop_flag && *stop_flag) return;
    const int nw = (int)gridDim.x * kWavesPerBlock;
    for (int t = (int)blockIdx.x * kWavesPerBlock + (int)(threadIdx.x >> 6); t < nslices; t += nw)
        solve_slice<T, V, BS, UPPER, true>(t, sinfo, blk_list, rw, next, scol, ve, vi, dinv, sx, x, y, in_scale,
                                           reset, err, copy, ep_a, ep_b, pre);
}

// The whole triangle in one launch of 16-wave workgroups, each solving 16 consecutive
// slices.  A workgroup draws a ticket when it starts; tickets follow arrival, and ticket
// q solves slices 16q .. 16q+15.  A gather only waits on rows of earlier levels, i.e. on
// smaller slice numbers, held by workgroups that drew smaller tickets and are running or
// done: the launch drains without assuming that all of it is resident.  The ticket
// counter is never reset: every launch advances it by exactly its workgroup count (a
// stopped launch in one step), and the host passes the count of earlier launches.
constexpr int kTicketWaves = 16;
template <typename T, typename V, int BS, bool UPPER>
__global__ __launch_bounds__(64 * kTicketWaves) void k_sptrsv_ticket(
    int nslices, unsigned base, const int4 *__restrict__ sinfo, const int *__restrict__ blk_list,
    const typename RowWord<BS>::W *__restrict__ rw, const int *__restrict__ next, const int *__restrict__ scol,
    const V *__restrict__ ve, const V *__restrict__ vi, const DiagT<V> *__restrict__ dinv,
    const signed char *__restrict__ sx, const T *x, T *y, const double *__restrict__ in_scale, T *reset,
    unsigned *sync, T *copy, const T *ep_a, T *ep_b, const int *__restrict__ stop_flag, SweepPre<T> pre) {
    if (stop_flag && *stop_flag) {
        if (blockIdx.x == 0 && threadIdx.x == 0)
            __hip_atomic_fetch_add(sync + kSweepTicket, gridDim.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    __shared__ unsigned ticket;
    if (threadIdx.x == 0)
        ticket = __hip_atomic_fetch_add(sync + kSweepTicket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - base;
    __syncthreads();
    const unsigned q = ticket;
    if (q >= gridDim.x) {  // launches out of step with the host's count
        if (threadIdx.x == 0) __hip_atomic_store(sync + kSweepErr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    const int slice = (int)q * kTicketWaves + (int)(threadIdx.x >> 6);
    if (slice < nslices)
        solve_slice<T, V, BS, UPPER, true>(slice, sinfo, blk_list, rw, next, scol, ve, vi, dinv, sx, x, y, in_scale,
                                           reset, sync + kSweepErr, copy, ep_a, ep_b, pre);
}

// Mode 3: one stream of tickets per XCD, and no counter.  Workgroup b belongs to stream
// x = b mod 8 (the XCD the dispatcher hands it to) and holds that stream's ticket
// j = b / 8; stream x solves the x-th eighth of every level (a band of the domain), so
// most gathers find their rows in the XCD's own L2.  Every stream has the same number of
// tickets per level (bands padded with empty slices), so ticket j of any stream is in the
// same level, and a wait is only on rows of tickets < j, i.e. of workgroups with smaller
// numbers.  Progress needs only that each dispatcher hands out its workgroups in
// increasing order: the smallest unfinished workgroup then has its dispatcher's
// resources (every earlier workgroup of that dispatcher is finished) and waits on nothing
// unfinished.  Any other order would end in the bounded poll's SENK_ERR_TIMEOUT, not a
// hang.  (Tickets drawn from one atomic counter per XCD, which need only the in-order
// dispatch across XCDs, cost an atomic round trip before each workgroup's first load:
// FP32 application at 128^3, simple blocks of 8, 46.6 against 40.3 us.)
// In ticket order (xslice null) the slices themselves are numbered so that wave w of
// ticket j of stream x solves slice 64 j + 8 x + w: the wave goes from its workgroup
// number to its slice descriptor and row words without a lookup.  Otherwise
// xslice[x][j][w] names the slice (-1: none).
template <typename T, typename V, int BS, bool UPPER, int CH, bool PIPE, bool O32>
__global__ __launch_bounds__(64 * kXTicketWaves) void k_sptrsv_xticket(
    int xwg, const int *__restrict__ xslice, const int4 *__restrict__ sinfo, const int *__restrict__ blk_list,
    const typename RowWord<BS>::W *__restrict__ rw, const int *__restrict__ next, const int *__restrict__ scol,
    const V *__restrict__ ve, const V *__restrict__ vi, const DiagT<V> *__restrict__ dinv,
    const signed char *__restrict__ sx, const T *x, T *y, const double *__restrict__ in_scale, T *reset,
    unsigned *sync, T *copy, const T *ep_a, T *ep_b, const int *__restrict__ stop_flag, SweepPre<T> pre) {
    const int xs = (int)(blockIdx.x % kXcds), q = (int)(blockIdx.x / kXcds);
    if (q >= xwg) return;
    const int wv = (int)(threadIdx.x >> 6);
    const int slice = xslice ? xslice[(size_t)xs * kXTicketWaves * xwg + (size_t)q * kXTicketWaves + wv]
                             : q * (kXcds * kXTicketWaves) + xs * kXTicketWaves + wv;
    if (slice < 0) return;
    // the stop flag is read together with the slice's head (both addresses are known)
    const int stopped = stop_flag ? *stop_flag : 0;
    const SliceHead<V, BS> h = slice_head<V, BS, UPPER, O32>(slice, sinfo, rw, next, dinv, sx);
    if (stopped) return;
    solve_slice_body<T, V, BS, UPPER, true, CH, PIPE, O32>(h, blk_list, scol, ve, vi, x, y, in_scale, reset,
                                                           sync + kSweepErr, copy, ep_a, ep_b, pre);
}

// The update of a U solve's right-hand side (SweepPre) as a pass of its own, for the
// level-by-level mode; the same arithmetic as the sweeps' prologue.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_sweep_pre(int n, SweepPre<T> pre, const int *__restrict__ stop) {
    if (stop && *stop) return;
    const int i = (int)(blockIdx.x * kBlock + threadIdx.x);
    if (i >= n) return;
    const T a = pre.ab[0], be = pre.ab[1];
    const T u = pre.u[i], r0 = pre.r[i], qv = pre.q[i], yv = pre.yd[i];
    const T rv = (-a) * qv + r0;
    pre.yd[i] = a * u + yv;
    pre.r[i] = rv;
    pre.u[i] = rv + be * u;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_fill_sentinel(int n, T *p) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = __builtin_bit_cast(T, Bits<T>::kSentinel);
}

// ----------------------------------------------------------------- host side
namespace {
inline int slices_of(int nblk, int bs) { return (int)(((long long)nblk * bs + 63) / 64); }
inline int waves_grid(int nsl) { return (nsl + kWavesPerBlock - 1) / kWavesPerBlock; }
bool layout_bsize_ok(int bs) { return bs >= 1 && bs <= 64 && (64 % bs) == 0; }

template <int BS>
int count_rows(Ctx *c, const Schedule &s, const TriLayout &L, const int *rptr, const int *cind, int *ecnt, int *icnt,
               int *err) {
    using W = typename RowWord<BS>::W;
    if (L.upper)
        k_tri_count<BS, true><<<waves_grid(L.nslices), kBlock, 0, c->stream>>>(
            L.nslices, L.sinfo, s.dblk, rptr, cind, (W *)L.rw, L.next, ecnt, icnt, err);
    else
        k_tri_count<BS, false><<<waves_grid(L.nslices), kBlock, 0, c->stream>>>(
            L.nslices, L.sinfo, s.dblk, rptr, cind, (W *)L.rw, L.next, ecnt, icnt, err);
    SENK_LAUNCH_CHECK();
    return SENK_OK;
}

template <typename S, typename V, int BS>
int fill_rows(Ctx *c, const Schedule &s, const TriLayout &L, const int *rptr, const int *cind, const S *val, int *scol,
              V *ve, V *vi, DiagT<V> *dinv, signed char *sx) {
    using W = typename RowWord<BS>::W;
    if (L.upper)
        k_tri_fill<S, V, BS, true><<<waves_grid(L.nslices), kBlock, 0, c->stream>>>(
            L.nslices, L.sinfo, s.dblk, rptr, cind, val, (const W *)L.rw, L.next, scol, ve, vi, dinv, sx);
    else
        k_tri_fill<S, V, BS, false><<<waves_grid(L.nslices), kBlock, 0, c->stream>>>(
            L.nslices, L.sinfo, s.dblk, rptr, cind, val, (const W *)L.rw, L.next, scol, ve, vi, dinv, sx);
    SENK_LAUNCH_CHECK();
    return SENK_OK;
}

template <typename T, typename V, int BS>
int solve_levels(Ctx *c, const Schedule &s, const TriLayout &L, const V *ve, const V *vi, const DiagT<V> *dinv,
                 const signed char *sx, const T *x, T *y, const double *in_scale, const int *stop) {
    using W = typename RowWord<BS>::W;
    for (int q = 0; q < L.nsteps; q++) {
        const int sl0 = L.step_slice[q], nsl = L.step_slice[q + 1] - sl0;
        if (L.upper)
            k_sptrsv_sell<T, V, BS, true><<<waves_grid(nsl), kBlock, 0, c->stream>>>(
                sl0, nsl, L.tord ? 1 : 0, L.sinfo, s.dblk, (const W *)L.rw, L.next, L.scol, ve, vi, dinv, sx, x, y,
                nullptr, stop);
        else
            k_sptrsv_sell<T, V, BS, false><<<waves_grid(nsl), kBlock, 0, c->stream>>>(
                sl0, nsl, L.tord ? 1 : 0, L.sinfo, s.dblk, (const W *)L.rw, L.next, L.scol, ve, vi, dinv, sx, x, y,
                in_scale, stop);
        SENK_LAUNCH_CHECK();
    }
    return SENK_OK;
}

// Cooperative grid of a sweep kernel on the current device: half of the workgroups
// that are resident at once by the runtime's occupancy and by a conservative count of
// our own (VGPRs rounded up to 16 of the 512 per SIMD lane, 4 SIMDs per CU, at most 8
// waves each).  Every wave must be resident for the sweep to drain; the margin keeps it
// so whatever the register allocation granule is.
template <typename T, typename V, int BS, bool UPPER>
int sweep_grid() {
    static int cached[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
    if (!cached[dev]) {
        int per_cu = 0, cus = 0;
        hipFuncAttributes fa{};
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_sptrsv_sweep<T, V, BS, UPPER>, kBlock, 0) !=
                hipSuccess ||
            hipFuncGetAttributes(&fa, (const void *)k_sptrsv_sweep<T, V, BS, UPPER>) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return 0;
        const int vgpr = std::max(16, (fa.numRegs + 15) / 16 * 16);
        const int waves_per_simd = std::min(8, 512 / vgpr);
        const int own = waves_per_simd * 4 / kWavesPerBlock;  // workgroups per CU
        if (fa.localSizeBytes > 0 || fa.sharedSizeBytes > 0) return 0;
        cached[dev] = std::max(1, std::min(per_cu, own) * cus / 2);
    }
    return cached[dev];
}

template <typename T, typename V, int BS, bool UPPER>
int launch_sweep(Ctx *c, const Schedule &s, const TriLayout &L, const V *ve, const V *vi, const DiagT<V> *dinv,
                 const signed char *sx, const T *x, T *y, const double *in_scale, T *reset, T *copy,
                 const T *ep_a, T *ep_b, const int *stop, SweepPre<T> pre) {
    using W = typename RowWord<BS>::W;
    int grid = sweep_grid<T, V, BS, UPPER>();
    if (grid <= 0) return SENK_ERR_HIP;
    grid = std::min(grid, waves_grid(L.nslices));
    int nslices = L.nslices;
    const int4 *sinfo = L.sinfo;
    const int *blk = s.dblk, *next = L.next, *scol = L.scol;
    const W *rw = (const W *)L.rw;
    unsigned *err = L.sync + kSweepErr;
    const double *sc = UPPER ? nullptr : in_scale;
    void *args[] = {&nslices, &sinfo, &blk,  &rw,   &next,  &scol,  &ve,    &vi,  &dinv, &sx,
                    &x,       &y,     &sc,   &reset, &err, &copy, &ep_a, &ep_b, &stop, &pre};
    SENK_HIP(hipLaunchCooperativeKernel((const void *)k_sptrsv_sweep<T, V, BS, UPPER>, dim3(grid), dim3(kBlock), args,
                                        0, c->stream));
    return SENK_OK;
}

template <typename T, typename V, int BS, bool UPPER>
int launch_ticket(Ctx *c, const Schedule &s, TriLayout &L, const V *ve, const V *vi, const DiagT<V> *dinv,
                  const signed char *sx, const T *x, T *y, const double *in_scale, T *reset, T *copy,
                  const T *ep_a, T *ep_b, const int *stop, const SweepPre<T> &pre) {
    using W = typename RowWord<BS>::W;
    const unsigned grid = (unsigned)((L.nslices + kTicketWaves - 1) / kTicketWaves);
    const unsigned base = L.ticket_base;
    k_sptrsv_ticket<T, V, BS, UPPER><<<grid, 64 * kTicketWaves, 0, c->stream>>>(
        L.nslices, base, L.sinfo, s.dblk, (const W *)L.rw, L.next, L.scol, ve, vi, dinv, sx, x, y,
        UPPER ? nullptr : in_scale, reset, L.sync, copy, ep_a, ep_b, stop, pre);
    SENK_LAUNCH_CHECK();
    L.ticket_base = base + grid;
    return SENK_OK;
}

template <typename T, typename V, int BS, bool UPPER>
int launch_xticket(Ctx *c, const Schedule &s, TriLayout &L, const V *ve, const V *vi, const DiagT<V> *dinv,
                   const signed char *sx, const T *x, T *y, const double *in_scale, T *reset, T *copy,
                   const T *ep_a, T *ep_b, const int *stop, const SweepPre<T> &pre) {
    using W = typename RowWord<BS>::W;
    if ((!L.xslice && !L.tord) || L.xwg <= 0) return SENK_ERR_ARG;
    // FP64 rows of at most 8 external entries: rounds of 4 gathers instead of one of 8
    // (64-66 VGPRs instead of 96-98: 7-8 waves per SIMD instead of 4-5; FP64 application at
    // 128^3, blocks of 8, 57.0 against 64.2 us in the simple order, 86.4 against 89.0 us in
    // the connect order, whose rows have up to 6).  FP32 (67 VGPRs either way) loses with
    // a second round: 72.3 against 62.9 us in the connect order.
    const bool ch4 = std::is_same<T, double>::value && L.max_ext <= 8;
    const bool pipe = L.max_ext > (ch4 ? 4 : 8);
    // 32-bit byte offsets where every array the sweep indexes fits them (elements of at most
    // 8 bytes; the slice descriptors, 16 bytes each, number at most a 64th of the slots)
    const long long most = std::max({(long long)L.n, L.n_ext, L.n_int, (long long)L.nslices * 64});
    const bool o32 = most * 8 < (1LL << 32);
#define SENK_XT_LAUNCH1(CHV, PIPEV, O32V)                                                                      \
    k_sptrsv_xticket<T, V, BS, UPPER, CHV, PIPEV, O32V><<<kXcds * L.xwg, 64 * kXTicketWaves, 0, c->stream>>>(  \
        L.xwg, L.xslice, L.sinfo, s.dblk, (const W *)L.rw, L.next, L.scol, ve, vi, dinv, sx, x, y,              \
        UPPER ? nullptr : in_scale, reset, L.sync, copy, ep_a, ep_b, stop, pre)
#define SENK_XT_LAUNCH(CHV, PIPEV)              \
    do {                                        \
        if (o32) SENK_XT_LAUNCH1(CHV, PIPEV, true);  \
        else SENK_XT_LAUNCH1(CHV, PIPEV, false); \
    } while (0)
    // rows of more than 32 external entries (an irregular matrix's long tail): rounds of 16,
    // at 108-161 VGPRs, as the longest row of a level sets its time (unstructured n = 2^22,
    // blocks of 16: FP32 / FP64 application 944 / 1048 us against 1150 / 1260 us with rounds
    // of 8; the 27-point system, rows of 17-32, loses with them: FP64 315 against 246 us)
    if (ch4 && pipe) SENK_XT_LAUNCH(4, true);
    else if (ch4) SENK_XT_LAUNCH(4, false);
    else if (L.max_ext > 32) SENK_XT_LAUNCH(16, true);
    else if (pipe) SENK_XT_LAUNCH(8, true);
    else SENK_XT_LAUNCH(8, false);
#undef SENK_XT_LAUNCH
#undef SENK_XT_LAUNCH1
    SENK_LAUNCH_CHECK();
    return SENK_OK;
}

template <typename T, typename V, int BS>
int solve_sweep(Ctx *c, const Schedule &s, TriLayout &L, const V *ve, const V *vi, const DiagT<V> *dinv,
                const signed char *sx, const T *x, T *y, const double *in_scale, T *reset, T *copy,
                const int *stop, const T *ep_a = nullptr, T *ep_b = nullptr,
                const SweepPre<T> &pre = SweepPre<T>{}) {
    const int mode = L.mode;
    if (mode == kTrsvXcd) {
        if (L.upper)
            return launch_xticket<T, V, BS, true>(c, s, L, ve, vi, dinv, sx, x, y, in_scale, reset, copy, ep_a,
                                                  ep_b, stop, pre);
        return launch_xticket<T, V, BS, false>(c, s, L, ve, vi, dinv, sx, x, y, in_scale, reset, copy, ep_a, ep_b,
                                               stop, SweepPre<T>{});
    }
    if (mode == kTrsvTicket) {
        if (L.upper)
            return launch_ticket<T, V, BS, true>(c, s, L, ve, vi, dinv, sx, x, y, in_scale, reset, copy, ep_a,
                                                 ep_b, stop, pre);
        return launch_ticket<T, V, BS, false>(c, s, L, ve, vi, dinv, sx, x, y, in_scale, reset, copy, ep_a, ep_b,
                                              stop, SweepPre<T>{});
    }
    if (L.upper)
        return launch_sweep<T, V, BS, true>(c, s, L, ve, vi, dinv, sx, x, y, in_scale, reset, copy, ep_a, ep_b,
                                            stop, pre);
    return launch_sweep<T, V, BS, false>(c, s, L, ve, vi, dinv, sx, x, y, in_scale, reset, copy, ep_a, ep_b, stop,
                                         SweepPre<T>{});
}

template <typename T>
int fill_sentinel(Ctx *c, int n, T *p) {
    if (n <= 0) return SENK_OK;
    k_fill_sentinel<T><<<(n + kBlock - 1) / kBlock, kBlock, 0, c->stream>>>(n, p);
    SENK_LAUNCH_CHECK();
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
size_t row_word_bytes(int bs) { return bs <= 16 ? sizeof(unsigned) : sizeof(unsigned long long); }
}  // namespace

void tri_layout_free(TriLayout *L) {
    for (void *p : {(void *)L->sinfo, (void *)L->next, L->rw, (void *)L->scol, (void *)L->ve64, (void *)L->vi64,
                    (void *)L->d64, (void *)L->ve32, (void *)L->vi32, (void *)L->d32, (void *)L->sync, L->mid64,
                    L->mid32, L->ve16, L->vi16, (void *)L->d16, (void *)L->sx16, (void *)L->xslice})
        if (p) (void)hipFree(p);
    *L = TriLayout{};
}

int tri_layout_build(Ctx *c, const Schedule &s, bool upper, int N, const int *rptr, const int *cind,
                     TriLayout *L) {
    const int BS = s.bsize;
    if (!layout_bsize_ok(BS) || (long long)s.lev_ptr.size() < (long long)s.nlev + 1) return SENK_ERR_UNSUPPORTED;
    if ((long long)s.lev_ptr[s.nlev] * BS != N) return SENK_ERR_SCHEDULE;
    tri_layout_free(L);
    L->bsize = BS;
    L->upper = upper;
    L->n = N;
    // steps = non-empty levels in solve order; slices numbered in that order.  Mode 3
    // streams: stream x takes the x-th eighth (band) of every step's slices, every step
    // padded to the same number of 8-slice tickets in every stream.
    const int per = 64 / BS, group = kXcds * kXTicketWaves;
    std::vector<int> lev_of, nsl_of, tk_of;
    long long plain = 0, padded = 0;
    for (int q = 0; q < s.nlev; q++) {
        const int l = upper ? s.nlev - 1 - q : q;
        const int nsl = slices_of(s.lev_ptr[l + 1] - s.lev_ptr[l], BS);
        if (nsl <= 0) continue;
        int tk = 0;
        for (int x = 0; x < kXcds; x++) {
            const int cnt = (int)((long long)nsl * (x + 1) / kXcds - (long long)nsl * x / kXcds);
            tk = std::max(tk, (cnt + kXTicketWaves - 1) / kXTicketWaves);
        }
        lev_of.push_back(l);
        nsl_of.push_back(nsl);
        tk_of.push_back(tk);
        plain += nsl;
        padded += (long long)tk * group;
    }
    // Ticket order (slice 64 j + 8 x + w = wave w of ticket j of stream x, empty slices
    // where a band runs short) when the padding at most doubles the slices; otherwise
    // slice order and a ticket -> slice table (many small levels, e.g. level schedules).
    L->tord = padded <= 2 * plain && padded <= INT32_MAX;
    std::vector<int4> si;
    std::vector<int> table[kXcds];
    L->step_slice.assign(1, 0);
    for (size_t q = 0; q < lev_of.size(); q++) {
        const int l = lev_of[q], nsl = nsl_of[q], tk = tk_of[q];
        const int b0 = s.lev_ptr[l], nb = s.lev_ptr[l + 1] - b0;
        auto slice_of = [&](int w) { return make_int4(0, 0, b0 + w * per, std::min(per, nb - w * per)); };
        if (L->tord) {
            for (int k = 0; k < tk; k++)
                for (int x = 0; x < kXcds; x++) {
                    const int lo = (int)((long long)nsl * x / kXcds), hi = (int)((long long)nsl * (x + 1) / kXcds);
                    for (int w = 0; w < kXTicketWaves; w++) {
                        const int ls = lo + k * kXTicketWaves + w;
                        si.push_back(ls < hi ? slice_of(ls) : make_int4(0, 0, 0, 0));
                    }
                }
        } else {
            const int a = (int)si.size();
            for (int w = 0; w < nsl; w++) si.push_back(slice_of(w));
            for (int x = 0; x < kXcds; x++) {
                const int lo = a + (int)((long long)nsl * x / kXcds), hi = a + (int)((long long)nsl * (x + 1) / kXcds);
                for (int k = 0; k < tk * kXTicketWaves; k++) table[x].push_back(lo + k < hi ? lo + k : -1);
            }
        }
        L->step_slice.push_back((int)si.size());
    }
    L->nsteps = (int)L->step_slice.size() - 1;
    L->nslices = (int)si.size();
    if (L->nslices == 0) return SENK_OK;
    L->xwg = (int)(padded / group);
    if (!L->tord) {
        std::vector<int> flat;
        flat.reserve((size_t)padded);
        for (int x = 0; x < kXcds; x++) flat.insert(flat.end(), table[x].begin(), table[x].end());
        if (hipMalloc(&L->xslice, flat.size() * sizeof(int)) != hipSuccess ||
            hipMemcpy(L->xslice, flat.data(), flat.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
            tri_layout_free(L);
            return SENK_ERR_ALLOC;
        }
    }
    const size_t nslots = (size_t)L->nslices * 64;
    int *ecnt = nullptr, *icnt = nullptr, *err = nullptr;
    int st = SENK_OK;
    std::vector<int> he(2 * (size_t)L->nslices), hi(L->nslices);
    int herr = 0;
    auto fail = [&](int code) {
        (void)hipFree(ecnt);
        (void)hipFree(icnt);
        (void)hipFree(err);
        if (code) tri_layout_free(L);
        return code;
    };
    if (hipMalloc(&L->sinfo, si.size() * sizeof(int4)) != hipSuccess ||
        hipMalloc(&L->rw, nslots * row_word_bytes(BS)) != hipSuccess ||
        (BS > 32 && hipMalloc(&L->next, nslots * sizeof(int)) != hipSuccess) ||
        hipMalloc(&ecnt, 2 * (size_t)L->nslices * sizeof(int)) != hipSuccess ||
        hipMalloc(&icnt, L->nslices * sizeof(int)) != hipSuccess || hipMalloc(&err, sizeof(int)) != hipSuccess ||
        hipMalloc(&L->sync, kSweepWords * sizeof(unsigned)) != hipSuccess)
        return fail(SENK_ERR_ALLOC);
    if (hipMemcpyAsync(L->sinfo, si.data(), si.size() * sizeof(int4), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipMemsetAsync(err, 0, sizeof(int), c->stream) != hipSuccess ||
        hipMemsetAsync(L->sync, 0, kSweepWords * sizeof(unsigned), c->stream) != hipSuccess)
        return fail(SENK_ERR_HIP);
    auto count = [&]() -> int { SENK_BS_DISPATCH(BS, count_rows<B>(c, s, *L, rptr, cind, ecnt, icnt, err)) };
    if ((st = count())) return fail(st);
    if (hipMemcpyAsync(he.data(), ecnt, he.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipMemcpyAsync(hi.data(), icnt, hi.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return fail(SENK_ERR_HIP);
    if (herr) return fail(SENK_ERR_UNSUPPORTED);
    long long ne = 0, ni = 0;
    int mx = 0;
    for (int i = 0; i < L->nslices; i++) {
        si[i].x = (int)ne;
        si[i].y = (int)ni;
        ne += he[i];
        ni += hi[i];
        mx = std::max(mx, he[(size_t)L->nslices + i]);
        if (ne > INT32_MAX || ni > INT32_MAX) return fail(SENK_ERR_UNSUPPORTED);
    }
    L->n_ext = ne;
    L->max_ext = mx;
    L->n_int = ni;
    if (hipMalloc(&L->scol, (size_t)(ne ? ne : 1) * sizeof(int)) != hipSuccess) return fail(SENK_ERR_ALLOC);
    if (hipMemcpyAsync(L->sinfo, si.data(), si.size() * sizeof(int4), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
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
    // each layout owns the coherent output vector of its sweep (L: the intermediate
    // vector, U: the final rows before they are copied out), kept all-sentinel between
    // applications
    void **mid = std::is_same<T, double>::value ? &L->mid64 : &L->mid32;
    if (!*mid) {
        SENK_HIP(hipMalloc(mid, (size_t)L->n * sizeof(T)));
        SENK_TRY(fill_sentinel<T>(c, L->n, (T *)*mid));
    }
    int *scol = L->cols_filled ? nullptr : L->scol;
    auto run = [&]() -> int {
        SENK_BS_DISPATCH(L->bsize, (fill_rows<T, T, B>(c, s, *L, rptr, cind, val, scol, *ve, *vi, *dv, nullptr)))
    };
    SENK_TRY(run());
    L->cols_filled = true;
    return SENK_OK;
}
template int tri_layout_fill<double>(Ctx *, const Schedule &, TriLayout *, const int *, const int *, const double *);
template int tri_layout_fill<float>(Ctx *, const Schedule &, TriLayout *, const int *, const int *, const float *);

int tri_layout_fill_f16(Ctx *c, const Schedule &s, TriLayout *L, const int *rptr, const int *cind, const double *val) {
    if (L->nslices == 0) return SENK_OK;
    if (!L->ve16) SENK_HIP(hipMalloc(&L->ve16, (size_t)(L->n_ext ? L->n_ext : 1) * sizeof(__half)));
    if (!L->vi16) SENK_HIP(hipMalloc(&L->vi16, (size_t)(L->n_int ? L->n_int : 1) * sizeof(__half)));
    if (!L->sx16) SENK_HIP(hipMalloc(&L->sx16, (size_t)L->nslices * 64));
    if (L->upper && !L->d16) SENK_HIP(hipMalloc(&L->d16, (size_t)L->nslices * 64 * sizeof(float)));
    if (!L->mid32) {
        SENK_HIP(hipMalloc(&L->mid32, (size_t)L->n * sizeof(float)));
        SENK_TRY(fill_sentinel<float>(c, L->n, (float *)L->mid32));
    }
    int *scol = L->cols_filled ? nullptr : L->scol;
    auto run = [&]() -> int {
        SENK_BS_DISPATCH(L->bsize, (fill_rows<double, __half, B>(c, s, *L, rptr, cind, val, scol, (__half *)L->ve16,
                                                                 (__half *)L->vi16, L->d16, L->sx16)))
    };
    SENK_TRY(run());
    L->cols_filled = true;
    return SENK_OK;
}

namespace {
template <typename T, typename V>
int ilu_layout_apply(Ctx *c, const Schedule &ls, TriLayout &lt, const Schedule &us, TriLayout &ut,
                     const V *lve, const V *lvi, const V *uve, const V *uvi, const DiagT<V> *udv,
                     const signed char *lsx, const signed char *usx, const T *in, T *out, const double *in_scale,
                     const int *stop) {
    T *mid = (T *)(std::is_same<T, double>::value ? lt.mid64 : lt.mid32);
    T *up = (T *)(std::is_same<T, double>::value ? ut.mid64 : ut.mid32);
    if (lt.mode != kTrsvLevels && mid && up && lt.sync && ut.sync) {
        // L: in -> mid, every row of up (the U sweep's coherent output) becomes the
        // sentinel; U: mid -> up, mid returns to the sentinel as its rows are read, and
        // each final row is also written to out with an ordinary (cached) store
        auto l = [&]() -> int {
            SENK_BS_DISPATCH(lt.bsize, (solve_sweep<T, V, B>(c, ls, lt, lve, lvi, nullptr, lsx, in, mid, in_scale, up,
                                                             (T *)nullptr, stop)))
        };
        auto u = [&]() -> int {
            SENK_BS_DISPATCH(ut.bsize,
                             (solve_sweep<T, V, B>(c, us, ut, uve, uvi, udv, usx, mid, up, nullptr, mid, out, stop)))
        };
        SENK_TRY(l());
        return u();
    }
    auto l = [&]() -> int {
        SENK_BS_DISPATCH(lt.bsize, (solve_levels<T, V, B>(c, ls, lt, lve, lvi, nullptr, lsx, in, out, in_scale, stop)))
    };
    auto u = [&]() -> int {
        SENK_BS_DISPATCH(ut.bsize, (solve_levels<T, V, B>(c, us, ut, uve, uvi, udv, usx, out, out, nullptr, stop)))
    };
    SENK_TRY(l());
    return u();
}
}  // namespace

template <typename T>
int launch_ilu_layout(Ctx *c, const Schedule &ls, TriLayout &lt, const Schedule &us, TriLayout &ut,
                      const T *in, T *out, const double *in_scale, const int *stop, bool f16) {
    if (lt.nslices == 0 || ut.nslices == 0) return lt.nslices == ut.nslices ? SENK_OK : SENK_ERR_ARG;
    if (lt.bsize != ls.bsize || ut.bsize != us.bsize || lt.upper || !ut.upper || lt.n != ut.n) return SENK_ERR_ARG;
    if constexpr (std::is_same<T, float>::value) {
        if (f16) {
            if (!lt.ve16 || !lt.vi16 || !ut.ve16 || !ut.vi16 || !ut.d16 || !lt.sx16 || !ut.sx16) return SENK_ERR_ARG;
            return ilu_layout_apply<float, __half>(c, ls, lt, us, ut, (const __half *)lt.ve16, (const __half *)lt.vi16,
                                                   (const __half *)ut.ve16, (const __half *)ut.vi16, ut.d16, lt.sx16,
                                                   ut.sx16, in, out, in_scale, stop);
        }
    } else {
        if (f16) return SENK_ERR_ARG;
    }
    const bool f64 = std::is_same<T, double>::value;
    const T *lve = (const T *)(f64 ? (void *)lt.ve64 : (void *)lt.ve32);
    const T *lvi = (const T *)(f64 ? (void *)lt.vi64 : (void *)lt.vi32);
    const T *uve = (const T *)(f64 ? (void *)ut.ve64 : (void *)ut.ve32);
    const T *uvi = (const T *)(f64 ? (void *)ut.vi64 : (void *)ut.vi32);
    const T *udv = (const T *)(f64 ? (void *)ut.d64 : (void *)ut.d32);
    if (!lve || !lvi || !uve || !uvi || !udv) return SENK_ERR_ARG;
    return ilu_layout_apply<T, T>(c, ls, lt, us, ut, lve, lvi, uve, uvi, udv, nullptr, nullptr, in, out, in_scale, stop);
}
// One triangle on its own, in = out allowed.  In the single-launch modes, given both
// epilogue vectors (*fused tells whether they were used): a U sweep also writes
// ep_b = in - ep_a * out; an L sweep writes out = ep_a * ep_b + L^-1 in (ep_b read, and
// then not the same vector as out).  A U solve given pre (pre.u == in) first updates its
// right-hand side: in the sweep where a single launch solves the triangle, else by
// k_sweep_pre before the levels.  The single-launch modes keep their
// invariant as long as the triangles of one factorization alternate (L, U, L, ...):
// each sweep writes its own coherent vector, which the other triangle's last sweep
// left all-sentinel, and returns the other triangle's vector to the sentinel.
template <typename T>
T *tri_single_sweep_vector(const TriLayout &self, const TriLayout &other) {
    const bool f64 = std::is_same<T, double>::value;
    T *mine = (T *)(f64 ? self.mid64 : self.mid32);
    const void *theirs = f64 ? other.mid64 : other.mid32;
    return self.mode != kTrsvLevels && mine && theirs && self.sync ? mine : nullptr;
}
template float *tri_single_sweep_vector<float>(const TriLayout &, const TriLayout &);
template double *tri_single_sweep_vector<double>(const TriLayout &, const TriLayout &);

template <typename T>
int launch_tri_single(Ctx *c, const Schedule &s, TriLayout &self, TriLayout &other, const T *in, T *out,
                      const int *stop, const T *ep_a, T *ep_b, bool *fused, const SweepPre<T> &pre) {
    if (fused) *fused = false;
    if (pre.u && (!self.upper || pre.u != in || !pre.r || !pre.q || !pre.yd || !pre.ab)) return SENK_ERR_ARG;
    if (self.nslices == 0) return SENK_OK;
    if (self.bsize != s.bsize || self.n != other.n || self.upper == other.upper) return SENK_ERR_ARG;
    const bool f64 = std::is_same<T, double>::value;
    const T *ve = (const T *)(f64 ? (void *)self.ve64 : (void *)self.ve32);
    const T *vi = (const T *)(f64 ? (void *)self.vi64 : (void *)self.vi32);
    const T *dv = self.upper ? (const T *)(f64 ? (void *)self.d64 : (void *)self.d32) : nullptr;
    if (!ve || !vi || (self.upper && !dv)) return SENK_ERR_ARG;
    T *mine = tri_single_sweep_vector<T>(self, other);
    T *theirs = (T *)(f64 ? other.mid64 : other.mid32);
    if (!mine && !out) return SENK_ERR_ARG;
    if (mine) {
        if (!ep_a || !ep_b) ep_a = ep_b = nullptr;
        if (fused) *fused = ep_b != nullptr;
        SENK_BS_DISPATCH(self.bsize, (solve_sweep<T, T, B>(c, s, self, ve, vi, dv, nullptr, in, mine, nullptr, theirs,
                                                           out, stop, ep_a, ep_b, pre)))
    }
    if (pre.u) {
        k_sweep_pre<T><<<(self.n + kBlock - 1) / kBlock, kBlock, 0, c->stream>>>(self.n, pre, stop);
        SENK_LAUNCH_CHECK();
    }
    SENK_BS_DISPATCH(self.bsize, (solve_levels<T, T, B>(c, s, self, ve, vi, dv, nullptr, in, out, nullptr, stop)))
}
template int launch_tri_single<double>(Ctx *, const Schedule &, TriLayout &, TriLayout &, const double *, double *,
                                       const int *, const double *, double *, bool *, const SweepPre<double> &);
template int launch_tri_single<float>(Ctx *, const Schedule &, TriLayout &, TriLayout &, const float *, float *,
                                      const int *, const float *, float *, bool *, const SweepPre<float> &);
template int launch_ilu_layout<double>(Ctx *, const Schedule &, TriLayout &, const Schedule &, TriLayout &,
                                       const double *, double *, const double *, const int *, bool);
template int launch_ilu_layout<float>(Ctx *, const Schedule &, TriLayout &, const Schedule &, TriLayout &,
                                      const float *, float *, const double *, const int *, bool);

int tri_layout_check(Ctx *c, TriLayout *L) {
    if (!L->sync) return SENK_OK;
    unsigned err = 0;
    SENK_HIP(hipMemcpyAsync(&err, L->sync + kSweepErr, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
    SENK_HIP(hipStreamSynchronize(c->stream));
    if (!err) return SENK_OK;
    // a wait that gave up leaves sentinels or stale rows in the intermediate vector:
    // start the next application from a clean state
    SENK_HIP(hipMemsetAsync(L->sync, 0, kSweepWords * sizeof(unsigned), c->stream));
    L->ticket_base = 0;
    if (L->mid64) SENK_TRY(fill_sentinel<double>(c, L->n, (double *)L->mid64));
    if (L->mid32) SENK_TRY(fill_sentinel<float>(c, L->n, (float *)L->mid32));
    SENK_HIP(hipStreamSynchronize(c->stream));
    return SENK_ERR_TIMEOUT;
}

}  // namespace senk_amd
