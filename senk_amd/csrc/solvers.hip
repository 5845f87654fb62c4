// Device-resident Krylov drivers: GMRES(m)/FGMRES(m) (senk_gmres.hpp), GCR(m)/FGCR(m)
// (senk_gcr.hpp), BiCGStab (senk_bicgstab.hpp) and FP32-inner / FP64-refinement
// mixed precision.
//
// Vectors never leave the device.  Scalars (Hessenberg, Givens, alpha/beta/omega) live
// in device memory.  Every inner product is a streaming pass that leaves one partial
// sum per workgroup; the kernel that needs the scalar adds those partials in a fixed
// order in its prologue (or a one-workgroup finalize kernel does), so results are
// deterministic and no kernel waits on another workgroup.  The stop test writes a
// device flag that every later kernel of the iteration checks on entry; the host polls
// that flag through pinned memory a few iterations behind the device, so the queue
// never drains while it waits.
//
// GMRES keeps V_k unnormalised and stores sig_k = 1/||.||: every consumer multiplies
// by sig_k on the fly (the preconditioner input, the Gram-Schmidt coefficients and the
// final combination), which removes the separate scaling pass of the reference's
// Scal(1/h, V_{j+1}).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <type_traits>
#include <vector>

#include "dev_common.hpp"
#include "solvers.hpp"

namespace senk_amd {

namespace {

// {seq, stop} as one 64-bit word to the host ring (system scope, no fence: the host
// reads nothing else from the slot); see Poller::dev_mark
__device__ __forceinline__ void write_mark(unsigned long long *mark, unsigned seq, int stop) {
    __hip_atomic_store(mark, ((unsigned long long)seq << 32) | (unsigned)(stop != 0), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_SYSTEM);
}

template <typename T> const T *mvals(const MatView &A);
template <> const double *mvals<double>(const MatView &A) { return A.v64; }
template <> const float *mvals<float>(const MatView &A) { return A.v32; }
template <typename T> const T *svals(const senk_sellcs &s);
template <> const double *svals<double>(const senk_sellcs &s) { return s.val64; }
template <> const float *svals<float>(const senk_sellcs &s) { return s.val32; }
template <typename T> const T *bvals(const MatView &A);
template <> const double *bvals<double>(const MatView &A) { return A.bv64; }
template <> const float *bvals<float>(const MatView &A) { return A.bv32; }

// Partial-sum slots of the context (each kRedMaxK x kRedMaxBlocks doubles).
enum { kSlotMgsA = 0, kSlotMgsB = 1, kSlotA = 2, kSlotB = 3, kSlotC = 4, kSlotRes = 5 };

inline int ew_grid(long long n) {
    long long b = (n + kBlock * 4LL - 1) / (kBlock * 4LL);
    if (b < 1) b = 1;
    if (b > 8192) b = 8192;
    return (int)b;
}

// ----------------------------------------------------------------- small kernels
template <typename T>
__global__ __launch_bounds__(kBlock) void k_scale_copy(int N, const T *__restrict__ in, T *__restrict__ out,
                                                       const double *__restrict__ dscale, const int *stop) {
    if (stop && *stop) return;
    const double sc = dscale ? *dscale : 1.0;
    SENK_FOR4(i0, N) {
        T v[4];
        get4(in, i0, N, v);
        if (dscale)
#pragma unroll
            for (int u = 0; u < 4; u++) v[u] = (T)((double)v[u] * sc);
        put4(out, i0, N, v);
    }
}

// Residual y = b - A x (x == nullptr means x = 0), G lanes per row, with the
// workgroup partials of ||y||^2.
template <typename T, int G>
__global__ __launch_bounds__(kBlock) void k_residual(int N, const int *__restrict__ rptr,
                                                     const int *__restrict__ cind, const T *__restrict__ val,
                                                     const T *__restrict__ x, const T *__restrict__ b,
                                                     T *__restrict__ y, double *partials) {
    const int rows_per_pass = (int)((long long)gridDim.x * kBlock / G);
    const long long gid = (long long)blockIdx.x * kBlock + threadIdx.x;
    const int q = (int)(gid % G);
    double ss = 0;
    for (int row = (int)(gid / G); row < N; row += rows_per_pass) {
        T acc = 0;
        if (x) {
            const int e = rptr[row + 1];
            for (int j = rptr[row] + q; j < e; j += G) acc += val[j] * x[cind[j]];
#pragma unroll
            for (int off = G / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, G);
        }
        if (q == 0) {
            // Axpby(1, b, -1, y): y = 1*b + (-1)*(A x)   (senk_gmres.hpp:215-216)
            const T r = (T)1 * b[row] + (T)(-1) * acc;
            y[row] = r;
            ss += (double)r * (double)r;
        }
    }
    const double v[1] = {ss};
    store_partials<1>(v, partials);
}

// nrm2 = sum of the partials, inv = 1/sqrt(nrm2)
__global__ __launch_bounds__(kBlock) void k_norm_fin(const double *P, int B, double *nrm2, double *inv) {
    const double s = sum_partials(P, B);
    if (threadIdx.x == 0) {
        *nrm2 = s;
        if (inv) *inv = 1.0 / sqrt(s);
    }
}

// y = 1*b + (-1)*y with the partial sums of y^2 (residual through the SELL or BCSR
// product: y holds A x on entry)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_sub_norm(int N, const T *__restrict__ b, T *y, double *partials) {
    double ss = 0;
    SENK_FOR4(i0, N) {
        T bv[4], yv[4];
        get4(b, i0, N, bv);
        get4(y, i0, N, yv);
#pragma unroll
        for (int u = 0; u < 4; u++) {
            yv[u] = (T)1 * bv[u] + (T)(-1) * yv[u];
            ss += (double)yv[u] * (double)yv[u];
        }
        put4(y, i0, N, yv);
    }
    const double v[1] = {ss};
    store_partials<1>(v, partials);
}

template <typename T>
int residual(Ctx *c, const MatView &A, const T *x, const T *b, T *y, int *grid_out) {
    constexpr int G = 8;
    const T *v = mvals<T>(A);
    if (x && A.sell && svals<T>(*A.sell)) {
        // the SELL product sums every row in its CSR order, as the reference's SpmvCsr does,
        // and streams the matrix at the SpMV's rate (128^3 FP64: 34 + 10 us against 112 us
        // for the CSR kernel with 8 lanes per row); then y = 1*b + (-1)*y
        SENK_TRY(spmv<T>(c, A, x, y, nullptr));
        const int grid = red_grid(A.N);
        k_sub_norm<T><<<grid, kBlock, 0, c->stream>>>(A.N, b, y, red_slot(c, kSlotRes));
        SENK_LAUNCH_CHECK();
        *grid_out = grid;
        return SENK_OK;
    }
    if (!v && A.bnl && bvals<T>(A)) {
        if (!x) return SENK_ERR_ARG;
        SENK_TRY(spmv<T>(c, A, x, y, nullptr));
        const int grid = red_grid(A.N);
        k_sub_norm<T><<<grid, kBlock, 0, c->stream>>>(A.N, b, y, red_slot(c, kSlotRes));
        SENK_LAUNCH_CHECK();
        *grid_out = grid;
        return SENK_OK;
    }
    if (!v) return SENK_ERR_ARG;
    long long threads = (long long)A.N * G;
    int grid = (int)((threads + kBlock - 1) / kBlock);
    if (grid > kRedMaxBlocks) grid = kRedMaxBlocks;
    k_residual<T, G><<<grid, kBlock, 0, c->stream>>>(A.N, A.rptr, A.cind, v, x, b, y, red_slot(c, kSlotRes));
    SENK_LAUNCH_CHECK();
    *grid_out = grid;
    return SENK_OK;
}

// ----------------------------------------------------------------- host-side stop polling
// The last kernel of an iteration writes {sequence number, stop flag} as one 64-bit word
// into a slot of the context's coherent pinned ring (no copy, no event per iteration);
// the host, some iterations behind, waits for each slot's number and reads the flag.
// Sequence numbers run on across solves (Ctx::dseq), so a slot's old value never passes
// for a new mark.
struct Poller {
    static constexpr int kRing = 8;
    Ctx *c = nullptr;
    long long issued = 0, checked = 0;
    bool stopped = false;
    unsigned expect[kRing] = {};
    int init(Ctx *cc, const int *) {
        c = cc;
        return ctx_events(c);
    }
    void reset() {
        issued = checked = 0;
        stopped = false;
    }
    // Slot for the next mark (device address) and its number; pass both to the kernel.
    unsigned long long *dev_mark(unsigned *seq) {
        const int slot = (int)(issued % kRing);
        *seq = ++c->dseq;
        expect[slot] = *seq;
        issued++;
        return reinterpret_cast<unsigned long long *>(c->dring_d + 4 * slot);
    }
    // Wait until at most `lag` marks are unchecked; returns stop seen.  A mark that has
    // not arrived when the stream has nothing left to run is an error, not a wait.
    int poll(int lag, bool *stop) {
        while (!stopped && issued - checked > lag) {
            const int slot = (int)(checked % kRing);
            const unsigned long long *w = reinterpret_cast<const unsigned long long *>(c->dring_h + 4 * slot);
            unsigned long long v;
            for (long long spin = 0; ((v = __atomic_load_n(w, __ATOMIC_ACQUIRE)) >> 32) != expect[slot]; spin++) {
                if ((spin & 4095) == 4095 && hipStreamQuery(c->stream) == hipSuccess &&
                    (__atomic_load_n(w, __ATOMIC_ACQUIRE) >> 32) != expect[slot]) {
                    record_error("solver: an iteration's mark did not arrive");
                    return SENK_ERR_HIP;
                }
            }
            if ((unsigned)v) stopped = true;
            checked++;
        }
        *stop = stopped;
        return SENK_OK;
    }
    // A ring slot is reused only after its mark was checked: callers poll with lag < kRing.
};
constexpr int kLag = 2;

// Device workspace carved from the context arena.
struct Carver {
    char *base;
    size_t off = 0, cap;
    template <typename T>
    T *take(size_t n) {
        off = (off + 255) & ~(size_t)255;
        T *p = (T *)(base + off);
        off += sizeof(T) * n;
        return p;
    }
};
template <typename T>
size_t bytes_of(size_t n) {
    return ((sizeof(T) * n + 255) & ~(size_t)255) + 256;
}

// Device time of a solve from two of the context's events; `outer` selects the pair
// kept for a loop that runs inner solves (which use the other pair).
struct SolveTimer {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int start(Ctx *c, bool outer = false) {
        SENK_TRY(ctx_events(c));
        e0 = c->ev[outer ? 10 : 8];
        e1 = c->ev[outer ? 11 : 9];
        SENK_HIP(hipEventRecord(e0, c->stream));
        return SENK_OK;
    }
    int stop(Ctx *c, float *ms) {
        SENK_HIP(hipEventRecord(e1, c->stream));
        SENK_HIP(hipEventSynchronize(e1));
        SENK_HIP(hipEventElapsedTime(ms, e0, e1));
        return SENK_OK;
    }
};

void monitor_learned(const PrecView &M, double rho, double rho_prev);

// ================================================================= GMRES kernels
// e[0] = ||r0||, sig[0] = 1/||r0||, state reset (cycle start).
__global__ __launch_bounds__(kBlock) void k_gmres_start(const double *P, int B, double *e, double *sig, int *state) {
    const double s = sum_partials(P, B);
    if (threadIdx.x == 0) {
        const double n = sqrt(s);
        e[0] = n;
        sig[0] = 1.0 / n;
        state[0] = 0;
        state[1] = 0;
        state[2] = 0;
    }
}

// One modified Gram-Schmidt step fused with the next inner product (k = 0 .. j+1):
//   prologue (k > 0): H[k-1] = sig[k-1] * sum(Pprev);  w -= (H[k-1]*sig[k-1]) V_{k-1}
//   then partials of (V_k, w) (k <= j) or of (w, w) (k == j+1) into Pout.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_mgs(int N, int k, int j, const T *__restrict__ V, size_t ld,
                                                const double *sig, double *hc, T *w, const double *Pprev, int B,
                                                double *Pout, const int *stop) {
    if (*stop) return;
    T coef = 0;
    if (k > 0) {
        const double h = sig[k - 1] * sum_partials(Pprev, B);
        if (blockIdx.x == 0 && threadIdx.x == 0) hc[k - 1] = h;
        coef = (T)(h * sig[k - 1]);
    }
    const T *Vprev = k > 0 ? V + (size_t)(k - 1) * ld : nullptr;
    const T *Vk = k <= j ? V + (size_t)k * ld : nullptr;
    double acc = 0;
    SENK_FOR4(i0, N) {
        T wv[4];
        get4(w, i0, N, wv);
        if (Vprev) {
            T pv[4];
            get4(Vprev, i0, N, pv);
#pragma unroll
            for (int u = 0; u < 4; u++) wv[u] = wv[u] - coef * pv[u];
            put4(w, i0, N, wv);
        }
        if (Vk) {
            T kv[4];
            get4(Vk, i0, N, kv);
#pragma unroll
            for (int u = 0; u < 4; u++) acc += (double)kv[u] * (double)wv[u];
        } else {
#pragma unroll
            for (int u = 0; u < 4; u++) acc += (double)wv[u] * (double)wv[u];
        }
    }
    const double v[1] = {acc};
    store_partials<1>(v, Pout);
}

// Classical Gram-Schmidt with re-orthogonalisation (CGS2), one pass over V_0..V_{K-1}
// (K = j+1 <= kRedMaxK):
//   coef != nullptr:  w -= sum_k coef[k]*sig[k] V_k  first
//   mode 0: partials of (V_k, w), k < K;   mode 1: partials of (w, w)
template <typename T, int KMAX>
__global__ __launch_bounds__(kBlock) void k_cgs(int N, int K, const T *__restrict__ V, size_t ld,
                                                const double *sig, const double *coef, T *w, int mode,
                                                double *Pout, const int *stop) {
    if (*stop) return;
    __shared__ double cf[KMAX];
    if (coef && threadIdx.x < K) cf[threadIdx.x] = coef[threadIdx.x] * sig[threadIdx.x];
    __syncthreads();
    double acc[KMAX];
#pragma unroll
    for (int k = 0; k < KMAX; k++) acc[k] = 0;
    SENK_FOR4(i0, N) {
        T wv[4];
        get4(w, i0, N, wv);
        if (coef) {
            double sv[4];
#pragma unroll
            for (int u = 0; u < 4; u++) sv[u] = (double)wv[u];
#pragma unroll
            for (int k = 0; k < KMAX; k++)
                if (k < K) {
                    T vk[4];
                    get4(V + (size_t)k * ld, i0, N, vk);
#pragma unroll
                    for (int u = 0; u < 4; u++) sv[u] -= cf[k] * (double)vk[u];
                }
#pragma unroll
            for (int u = 0; u < 4; u++) wv[u] = (T)sv[u];
            put4(w, i0, N, wv);
        }
        if (mode == 0) {
#pragma unroll
            for (int k = 0; k < KMAX; k++)
                if (k < K) {
                    T vk[4];
                    get4(V + (size_t)k * ld, i0, N, vk);
#pragma unroll
                    for (int u = 0; u < 4; u++) acc[k] += (double)vk[u] * (double)wv[u];
                }
        } else {
#pragma unroll
            for (int u = 0; u < 4; u++) acc[0] += (double)wv[u] * (double)wv[u];
        }
    }
    store_partials_n<KMAX>(acc, mode == 0 ? K : 1, Pout);
}
// k_cgs with its accumulator count sized to K
template <typename T>
void launch_cgs(Ctx *c, int B, int N, int K, const T *V, size_t ld, const double *sig, const double *coef, T *w,
                int mode, double *Pout, const int *stop) {
    if (K <= 8)
        k_cgs<T, 8><<<B, kBlock, 0, c->stream>>>(N, K, V, ld, sig, coef, w, mode, Pout, stop);
    else if (K <= 16)
        k_cgs<T, 16><<<B, kBlock, 0, c->stream>>>(N, K, V, ld, sig, coef, w, mode, Pout, stop);
    else if (K <= 24)
        k_cgs<T, 24><<<B, kBlock, 0, c->stream>>>(N, K, V, ld, sig, coef, w, mode, Pout, stop);
    else if (K <= 32)
        k_cgs<T, 32><<<B, kBlock, 0, c->stream>>>(N, K, V, ld, sig, coef, w, mode, Pout, stop);
    else
        k_cgs<T, kRedMaxK><<<B, kBlock, 0, c->stream>>>(N, K, V, ld, sig, coef, w, mode, Pout, stop);
}
// out[k] = sig[k] * sum(P_k), one workgroup per k
__global__ __launch_bounds__(kBlock) void k_cgs_fin(const double *P, int B, const double *sig, double *out,
                                                    const int *stop) {
    if (*stop) return;
    const int k = blockIdx.x;
    const double s = sum_partials(P + (size_t)k * kPartLd, B);
    if (threadIdx.x == 0) out[k] = sig[k] * s;
}

// Givens update of Hessenberg column j and the stop test (senk_gmres.hpp:228-246),
// FP64.  H[j+1,j] = ||w|| comes from the partials of the last Gram-Schmidt pass;
// with CGS2 the second projection's coefficients are added to the column first.
__global__ __launch_bounds__(kBlock) void k_gmres_tail(int j, int m, double *H, double *cs, double *sn, double *e,
                                                       double *sig, const double *Pn, int B, const double *add_h,
                                                       double tol, double nrm_b, int *state, double *hist, int hidx,
                                                       int hcap, unsigned long long *mark, unsigned seq) {
    if (state[0]) {
        if (mark && threadIdx.x == 0) write_mark(mark, seq, 1);
        return;
    }
    const double nn = sum_partials(Pn, B);
    if (threadIdx.x != 0) return;
    double *hc = H + (size_t)j * (m + 1);
    if (add_h)
        for (int k = 0; k <= j; k++) hc[k] += add_h[k];
    const double hn = sqrt(nn);
    hc[j + 1] = hn;
    sig[j + 1] = 1.0 / hn;
    for (int k = 0; k < j; k++) {
        const double a = hc[k], b = hc[k + 1];
        hc[k] = cs[k] * a - sn[k] * b;
        hc[k + 1] = sn[k] * a + cs[k] * b;
    }
    const double a = hc[j], b = hc[j + 1];
    const double r = sqrt(a * a + b * b);
    cs[j] = a / r;
    sn[j] = -b / r;
    hc[j] = r;
    hc[j + 1] = 0;
    e[j + 1] = sn[j] * e[j];
    e[j] = cs[j] * e[j];
    const double ae = fabs(e[j + 1]);
    if (hist && hidx < hcap) hist[hidx] = ae / nrm_b;
    state[1] = j + 1;
    if (!isfinite(ae) || !isfinite(r)) {
        state[2] = 1;
        state[0] = 1;
    } else if (ae <= tol) {
        state[0] = 1;
    }
    if (mark) write_mark(mark, seq, state[0]);
}

// Cycle end: y = Trsv(H, e) over the jdone columns (senk_blas2.hpp:24-34) and the
// combination coefficients y_k * sig_k (or y_k for FGMRES's Z vectors).
__global__ void k_gmres_y(int m, const double *H, const double *e, const double *sig, double *y, double *coef,
                          const int *state, int use_sig) {
    if (threadIdx.x != 0) return;
    const int J = state[1];
    for (int i = J - 1; i >= 0; i--) {
        double t = e[i];
        for (int jj = J - 1; jj > i; jj--) t -= H[(size_t)jj * (m + 1) + i] * y[jj];
        y[i] = t / H[(size_t)i * (m + 1) + i];
    }
    for (int k = 0; k < J; k++) coef[k] = use_sig ? y[k] * sig[k] : y[k];
}

// out = sum_{k<J} coef_k V_k (accumulate: out += ..., in the reference's sequential
// Axpy order).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_combine(int N, const T *__restrict__ V, size_t ld,
                                                    const double *__restrict__ coef, const int *state, T *out,
                                                    int accumulate) {
    const int J = state[1];
    SENK_FOR4(i0, N) {
        double acc[4];
        T ov[4];
        get4(out, i0, N, ov);
        int k0 = 0;
        if (!accumulate && J > 0) {
            T vk[4];
            get4(V, i0, N, vk);
#pragma unroll
            for (int u = 0; u < 4; u++) acc[u] = coef[0] * (double)vk[u];
            k0 = 1;
        } else {
#pragma unroll
            for (int u = 0; u < 4; u++) acc[u] = accumulate ? (double)ov[u] : 0.0;
        }
        for (int k = k0; k < J; k++) {
            T vk[4];
            get4(V + (size_t)k * ld, i0, N, vk);
#pragma unroll
            for (int u = 0; u < 4; u++) acc[u] += coef[k] * (double)vk[u];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) ov[u] = (T)acc[u];
        put4(out, i0, N, ov);
    }
}

}  // namespace

// ================================================================= dispatch
template <typename T>
int spmv(Ctx *c, const MatView &A, const T *x, T *y, const int *stop) {
    if (A.sell) {
        const T *sv = svals<T>(*A.sell);
        if (sv)
            return launch_spmv_sell<T>(c, A.N, A.sell->C, A.sell->nslices, A.sell->slice_ptr, A.sell->cind, sv,
                                       A.sell->perm, x, y, false, stop,
                                       (double)A.sell->nnz_padded / ((double)A.sell->nslices * A.sell->C),
                                       std::is_same<T, double>::value ? A.sell_order64 : A.sell_order32,
                                       std::is_same<T, double>::value ? A.sell_nt64 : A.sell_nt32,
                                       A.sell->nnz_padded);
    }
    if (A.bnl) {
        const T *bv = bvals<T>(A);
        if (bv) return launch_spmv_bcsr<T>(c, A.N, A.bnl, A.bnw, A.brptr, A.bcind, bv, x, y);
    }
    const T *v = mvals<T>(A);
    if (!v) return SENK_ERR_ARG;
    return launch_spmv_csr<T>(c, A.N, A.rptr, A.cind, v, x, y, A.avg_row, stop);
}
template int spmv<double>(Ctx *, const MatView &, const double *, double *, const int *);
template int spmv<float>(Ctx *, const MatView &, const float *, float *, const int *);

int learned_apply_f32(Ctx *c, const PrecView &M, const float *r_in, const double *in_scale, float *out,
                      const int *stop);
int learned_apply_f64(Ctx *c, const PrecView &M, const double *r_in, const double *in_scale, double *out,
                      const int *stop);

template <typename T>
int learned_apply(Ctx *c, const PrecView &M, const T *in, const double *in_scale, T *out, const int *stop);
template <>
int learned_apply<float>(Ctx *c, const PrecView &M, const float *in, const double *s, float *out, const int *stop) {
    return learned_apply_f32(c, M, in, s, out, stop);
}
template <>
int learned_apply<double>(Ctx *c, const PrecView &M, const double *in, const double *s, double *out,
                          const int *stop) {
    return learned_apply_f64(c, M, in, s, out, stop);
}

template <typename T>
static bool layout_has(const TriLayout *L) {
    if (!L) return false;
    return std::is_same<T, double>::value ? L->ve64 != nullptr : L->ve32 != nullptr;
}

template <typename T>
int prec_apply(Ctx *c, const PrecView &M, const T *in, const double *in_scale, T *out, T *tmp, const int *stop) {
    const int N = M.N;
    if constexpr (std::is_same<T, double>::value) {
        if (M.apply32 && M.kind != 0) {  // FP64 solve, FP32 factor application
            if (!M.in32 || !M.out32 || M.learned) return SENK_ERR_ARG;
            PrecView m32 = M;
            m32.apply32 = false;
            SENK_TRY(launch_convert<double, float>(c, N, in, M.in32, in_scale));
            SENK_TRY(prec_apply<float>(c, m32, M.in32, nullptr, M.out32, nullptr, stop));
            return launch_convert<float, double>(c, N, M.out32, out, nullptr);
        }
    }
    if (M.kind == 0) {
        k_scale_copy<T><<<ew_grid(N), kBlock, 0, c->stream>>>(N, in, out, in_scale, stop);
        SENK_LAUNCH_CHECK();
    } else if (M.kind == 1) {
        const T *lv = std::is_same<T, double>::value ? (const T *)M.l64 : (const T *)M.l32;
        const T *uv = std::is_same<T, double>::value ? (const T *)M.u64 : (const T *)M.u32;
        if (!lv || !uv) return SENK_ERR_ARG;
        if (std::is_same<T, float>::value && M.f16) {
            if (!M.lt || !M.ut) return SENK_ERR_ARG;
            SENK_TRY(launch_ilu_layout<T>(c, M.lts, *M.lt, M.uts, *M.ut, in, out, in_scale, stop, true));
        } else if (layout_has<T>(M.lt) && layout_has<T>(M.ut)) {
            SENK_TRY(launch_ilu_layout<T>(c, M.lts, *M.lt, M.uts, *M.ut, in, out, in_scale, stop));
        } else {
            SENK_TRY(launch_sptrsv<T>(c, M.ls, false, M.lrptr, M.lcind, lv, in, out, in_scale, stop));
            SENK_TRY(launch_sptrsv<T>(c, M.us, true, M.urptr, M.ucind, uv, out, out, nullptr, stop));
        }
    } else if (M.kind == 2) {
        const T *lv = std::is_same<T, double>::value ? (const T *)M.bl64 : (const T *)M.bl32;
        const T *uv = std::is_same<T, double>::value ? (const T *)M.bu64 : (const T *)M.bu32;
        if (!lv || !uv) return SENK_ERR_ARG;
        const T *src = in;
        if (in_scale) {
            if (!tmp) return SENK_ERR_ARG;
            k_scale_copy<T><<<ew_grid(N), kBlock, 0, c->stream>>>(N, in, tmp, in_scale, stop);
            SENK_LAUNCH_CHECK();
            src = tmp;
        }
        SENK_TRY(launch_sptrsv_bcsr<T>(c, M.ls, false, M.bnl, M.bnw, M.blrptr, M.blcind, lv, src, out));
        SENK_TRY(launch_sptrsv_bcsr<T>(c, M.us, true, M.bnl, M.bnw, M.burptr, M.bucind, uv, out, out));
    } else {
        return SENK_ERR_ARG;
    }
    if (M.learned) SENK_TRY(learned_apply<T>(c, M, in, in_scale, out, stop));
    return SENK_OK;
}
template int prec_apply<double>(Ctx *, const PrecView &, const double *, const double *, double *, double *,
                                const int *);
template int prec_apply<float>(Ctx *, const PrecView &, const float *, const double *, float *, float *,
                               const int *);

// ================================================================= GMRES / FGMRES
template <typename T>
static int gmres(Ctx *c, const MatView &A, const PrecView &M, const T *b, T *x, const SolveParams &p,
                 senk_result *res, double *hist, int hist_cap) {
    const int N = A.N, m = p.m;
    const size_t ld = pad4(N);  // basis stride: every column 16-byte aligned
    const bool flexible = p.solver == SENK_FGMRES;
    const int hcap = p.outer * m;
    size_t need = bytes_of<T>((size_t)(m + 1) * ld) + (flexible ? bytes_of<T>((size_t)m * ld) : 0) +
                  3 * bytes_of<T>(N) + bytes_of<double>((size_t)(m + 1) * m) + 7 * bytes_of<double>(m + 2) +
                  bytes_of<double>(hcap + 1) + bytes_of<int>(8);
    void *wsp;
    SENK_TRY(ctx_work(c, need, &wsp));
    Carver cv{(char *)wsp, 0, need};
    T *V = cv.take<T>((size_t)(m + 1) * ld);
    T *Z = flexible ? cv.take<T>((size_t)m * ld) : nullptr;
    T *t = cv.take<T>(N), *u = cv.take<T>(N), *tmp = cv.take<T>(N);
    double *H = cv.take<double>((size_t)(m + 1) * m);
    double *cs = cv.take<double>(m + 2), *sn = cv.take<double>(m + 2), *e = cv.take<double>(m + 2);
    double *y = cv.take<double>(m + 2), *coef = cv.take<double>(m + 2), *sig = cv.take<double>(m + 2);
    double *h2 = cv.take<double>(m + 2);
    double *dhist = cv.take<double>(hcap + 1);
    int *state = cv.take<int>(8);

    Poller poll;
    SENK_TRY(poll.init(c, state));
    SolveTimer tm;
    SENK_TRY(tm.start(c));

    const double tol = p.nrm_b * p.eps;
    const bool prec = M.kind != 0 || M.learned;
    const bool use_cgs = p.ortho == SENK_CGS2;
    // at most 512 workgroups in this driver's passes: each Gram-Schmidt step's consumer sums
    // every partial, and with 1024 of them the FP64 GMRES(30) at 128^3 took 125.9 ms, with
    // 512 121.8 ms (256: 132.3)
    const int B = std::min(red_grid(N), 512);
    int total = p.outer * m, converged = 0, breakdown = 0, cycles = 0;
    int hstate[4] = {0, 0, 0, 0};
    for (int i = 0; i < p.outer; i++) {
        cycles++;
        int rg = 0;
        SENK_TRY(residual<T>(c, A, x, b, V, &rg));
        k_gmres_start<<<1, kBlock, 0, c->stream>>>(red_slot(c, kSlotRes), rg, e, sig, state);
        SENK_LAUNCH_CHECK();
        poll.reset();
        bool stop = false;
        for (int j = 0; j < m && !stop; j++) {
            T *vj = V + (size_t)j * ld, *w = V + (size_t)(j + 1) * ld;
            T *z = flexible ? Z + (size_t)j * ld : t;
            // z = M^-1 (sig_j V_j);  w = A z
            SENK_TRY(prec_apply<T>(c, M, vj, sig + j, z, tmp, state));
            SENK_TRY(spmv<T>(c, A, z, w, state));
            double *hc = H + (size_t)j * (m + 1);
            const double *Pnorm;
            const double *add_h = nullptr;
            if (use_cgs && j + 1 <= kRedMaxK) {
                const int K = j + 1;
                // h1 = V^T w;  w -= V h1, h2 = V^T w;  w -= V h2, ||w||;  H[:,j] = h1 + h2
                launch_cgs<T>(c, B, N, K, V, ld, sig, nullptr, w, 0, red_slot(c, kSlotA), state);
                SENK_LAUNCH_CHECK();
                k_cgs_fin<<<K, kBlock, 0, c->stream>>>(red_slot(c, kSlotA), B, sig, hc, state);
                SENK_LAUNCH_CHECK();
                launch_cgs<T>(c, B, N, K, V, ld, sig, hc, w, 0, red_slot(c, kSlotB), state);
                SENK_LAUNCH_CHECK();
                k_cgs_fin<<<K, kBlock, 0, c->stream>>>(red_slot(c, kSlotB), B, sig, h2, state);
                SENK_LAUNCH_CHECK();
                launch_cgs<T>(c, B, N, K, V, ld, sig, h2, w, 1, red_slot(c, kSlotC), state);
                SENK_LAUNCH_CHECK();
                Pnorm = red_slot(c, kSlotC);
                add_h = h2;
            } else {
                for (int k = 0; k <= j + 1; k++) {
                    const double *Pin = red_slot(c, (k + 1) % 2 ? kSlotMgsB : kSlotMgsA);
                    double *Pout = red_slot(c, k % 2 ? kSlotMgsB : kSlotMgsA);
                    k_mgs<T><<<B, kBlock, 0, c->stream>>>(N, k, j, V, ld, sig, hc, w, Pin, B, Pout, state);
                    SENK_LAUNCH_CHECK();
                }
                Pnorm = red_slot(c, (j + 1) % 2 ? kSlotMgsB : kSlotMgsA);
            }
            unsigned seq = 0;
            unsigned long long *mark = poll.dev_mark(&seq);
            k_gmres_tail<<<1, kBlock, 0, c->stream>>>(j, m, H, cs, sn, e, sig, Pnorm, B, add_h, tol, p.nrm_b, state,
                                                      dhist, i * m + j, hcap, mark, seq);
            SENK_LAUNCH_CHECK();
            SENK_TRY(poll.poll(kLag, &stop));
        }
        // cycle end: x += M^-1 (V y)   (senk_gmres.hpp:248-255), or x += V y / Z y
        k_gmres_y<<<1, 64, 0, c->stream>>>(m, H, e, sig, y, coef, state, flexible ? 0 : 1);
        SENK_LAUNCH_CHECK();
        if (flexible) {
            k_combine<T><<<ew_grid(N), kBlock, 0, c->stream>>>(N, Z, ld, coef, state, x, 1);
            SENK_LAUNCH_CHECK();
        } else if (!prec) {
            k_combine<T><<<ew_grid(N), kBlock, 0, c->stream>>>(N, V, ld, coef, state, x, 1);
            SENK_LAUNCH_CHECK();
        } else {
            k_combine<T><<<ew_grid(N), kBlock, 0, c->stream>>>(N, V, ld, coef, state, u, 0);
            SENK_LAUNCH_CHECK();
            SENK_TRY(prec_apply<T>(c, M, u, nullptr, t, tmp, nullptr));
            SENK_TRY(launch_axpby<T>(c, N, 1.0, t, 1.0, x));
        }
        SENK_HIP(hipMemcpyAsync(hstate, state, sizeof(int) * 4, hipMemcpyDeviceToHost, c->stream));
        SENK_HIP(hipStreamSynchronize(c->stream));
        if (hstate[0]) {
            breakdown = hstate[2];
            if (!breakdown) {
                converged = 1;
                total = i * m + hstate[1];
            }
            break;
        }
    }
    float ms = 0;
    SENK_TRY(tm.stop(c, &ms));
    const int nh = std::min(hcap, converged ? total : cycles * m);
    std::vector<double> hh((size_t)std::max(nh, 1));
    if (nh > 0) SENK_HIP(hipMemcpy(hh.data(), dhist, sizeof(double) * nh, hipMemcpyDeviceToHost));
    if (hist)
        for (int k = 0; k < std::min(nh, hist_cap); k++) hist[k] = hh[k];
    res->iters = converged ? total : p.outer * m;
    res->converged = converged;
    res->restarts = cycles;
    res->implicit_relres = nh > 0 ? hh[nh - 1] : 0;
    res->t_solve_ms = ms;
    res->status = breakdown ? SENK_ERR_BREAKDOWN : SENK_OK;
    if (p.verbose) {
        if (converged) {
            std::printf("I iter %d\n", res->iters);
            std::printf("# res %e\n", res->implicit_relres);
        } else {
            std::printf("# iter %d\n", p.outer * m);
            std::printf("# res : Check by using senk::test\n");
        }
        std::fflush(stdout);
    }
    return SENK_OK;
}

// ================================================================= GCR / FGCR kernels
namespace {
// partials of (Ap_j, Ap_j) and (Ap_j, r)                        (senk_gcr.hpp:119-120)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gcr_alpha(int N, const T *__restrict__ Apj, const T *__restrict__ r,
                                                      double *Pout, const int *stop) {
    if (*stop) return;
    double a0 = 0, a1 = 0;
    SENK_FOR4(i0, N) {
        T av[4], rv[4];
        get4(Apj, i0, N, av);
        get4(r, i0, N, rv);
#pragma unroll
        for (int u = 0; u < 4; u++) {
            a0 += (double)av[u] * (double)av[u];
            a1 += (double)av[u] * (double)rv[u];
        }
    }
    const double v[2] = {a0, a1};
    store_partials<2>(v, Pout);
}
// alpha = (Ap_j, r)/(Ap_j, Ap_j); x += alpha p_j; r -= alpha Ap_j; partials of ||r||^2.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gcr_xr(int N, const T *__restrict__ pj, const T *__restrict__ Apj,
                                                   const double *Pin, int B, double *dotApj, T *x, T *r, double *Pout,
                                                   const int *stop) {
    if (*stop) return;
    const double d0 = sum_partials(Pin, B), d1 = sum_partials(Pin + kPartLd, B);
    if (blockIdx.x == 0 && threadIdx.x == 0) *dotApj = d0;
    const T a = (T)(d1 / d0);
    double ss = 0;
    SENK_FOR4(i0, N) {
        T xv[4], pv[4], rv[4], av[4];
        get4(x, i0, N, xv);
        get4(pj, i0, N, pv);
        get4(r, i0, N, rv);
        get4(Apj, i0, N, av);
#pragma unroll
        for (int u = 0; u < 4; u++) {
            xv[u] += a * pv[u];
            rv[u] = rv[u] + (-a) * av[u];
            ss += (double)rv[u] * (double)rv[u];
        }
        put4(x, i0, N, xv);
        put4(r, i0, N, rv);
    }
    const double v[1] = {ss};
    store_partials<1>(v, Pout);
}
// per-step stop test on ||r|| (check_every mode)
__global__ __launch_bounds__(kBlock) void k_gcr_check(const double *Pin, int B, double tol, double nrm_b, int hidx,
                                                      int jstep, int *state, double *hist, int hcap) {
    if (state[0]) return;
    const double nr = sqrt(sum_partials(Pin, B));
    if (threadIdx.x != 0) return;
    if (hist && hidx < hcap) hist[hidx] = nr / nrm_b;
    state[1] = jstep + 1;
    if (!isfinite(nr)) {
        state[2] = 1;
        state[0] = 1;
    } else if (nr < tol) {
        state[0] = 1;
    }
}
// partials of (Ap_k, AKr), k <= j                                 (senk_gcr.hpp:126-133)
template <typename T, int KMAX>
__global__ __launch_bounds__(kBlock) void k_gcr_betas(int N, int K, const T *__restrict__ Ap, size_t ld,
                                                      const T *__restrict__ AKr, double *Pout, const int *stop) {
    if (*stop) return;
    double acc[KMAX];
#pragma unroll
    for (int k = 0; k < KMAX; k++) acc[k] = 0;
    SENK_FOR4(i0, N) {
        T a[4];
        get4(AKr, i0, N, a);
#pragma unroll
        for (int k = 0; k < KMAX; k++)
            if (k < K) {
                T v[4];
                get4(Ap + (size_t)k * ld, i0, N, v);
#pragma unroll
                for (int u = 0; u < 4; u++) acc[k] += (double)v[u] * (double)a[u];
            }
    }
    store_partials_n<KMAX>(acc, K, Pout);
}
// beta_k = -(Ap_k, AKr) / dot_Ap[k], one workgroup per k
__global__ __launch_bounds__(kBlock) void k_gcr_beta_fin(const double *P, int B, const double *dotAp, double *beta,
                                                         const int *stop) {
    if (*stop) return;
    const int k = blockIdx.x;
    const double s = sum_partials(P + (size_t)k * kPartLd, B);
    if (threadIdx.x == 0) beta[k] = -s / dotAp[k];
}
// p_{j+1} = beta_0 p_0 + Kr + sum_{k>=1} beta_k p_k, and the same for Ap  (:127-133).
// Also stores the partials of (Ap_{j+1}, Ap_{j+1}) and (Ap_{j+1}, r) that the next
// step's alpha needs (k_gcr_alpha's output), so that pass is not repeated.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gcr_newdir(int N, int K, T *p, T *Ap, size_t ld, const T *Kr,
                                                       const T *AKr, const double *beta, const T *__restrict__ r,
                                                       double *Palpha, const int *stop, unsigned long long *mark,
                                                       unsigned seq) {
    // the step's stop test ran before: its flag is this step's mark
    if (mark && blockIdx.x == 0 && threadIdx.x == 0) write_mark(mark, seq, *stop);
    if (*stop) return;
    double a0 = 0, a1 = 0;
    SENK_FOR4(i0, N) {
        T a[4], b[4], t[4];
        get4(p, i0, N, a);
        get4(Kr, i0, N, t);
        const T b0 = (T)beta[0];
#pragma unroll
        for (int u = 0; u < 4; u++) a[u] = b0 * a[u] + t[u];
        get4(Ap, i0, N, b);
        get4(AKr, i0, N, t);
#pragma unroll
        for (int u = 0; u < 4; u++) b[u] = b0 * b[u] + t[u];
        for (int k = 1; k < K; k++) {
            const T bk = (T)beta[k];
            get4(p + (size_t)k * ld, i0, N, t);
#pragma unroll
            for (int u = 0; u < 4; u++) a[u] += bk * t[u];
            get4(Ap + (size_t)k * ld, i0, N, t);
#pragma unroll
            for (int u = 0; u < 4; u++) b[u] += bk * t[u];
        }
        put4(p + (size_t)K * ld, i0, N, a);
        put4(Ap + (size_t)K * ld, i0, N, b);
        get4(r, i0, N, t);
#pragma unroll
        for (int u = 0; u < 4; u++) {
            a0 += (double)b[u] * (double)b[u];
            a1 += (double)b[u] * (double)t[u];
        }
    }
    const double v[2] = {a0, a1};
    store_partials<2>(v, Palpha);
}
__global__ __launch_bounds__(kBlock) void k_gcr_restart(const double *P, int B, double tol, double nrm_b, int *state,
                                                        double *hist, int hidx, int hcap) {
    const double nr = sqrt(sum_partials(P, B));
    if (threadIdx.x != 0) return;
    if (hist && hidx < hcap) hist[hidx] = nr / nrm_b;
    state[0] = 0;
    state[1] = 0;
    state[2] = 0;
    if (!isfinite(nr)) {
        state[2] = 1;
        state[0] = 1;
    } else if (nr < tol) {
        state[0] = 1;
    }
}
}  // namespace

template <typename T>
static int gcr(Ctx *c, const MatView &A, const PrecView &M, const T *b, T *x, const SolveParams &p,
               senk_result *res, double *hist, int hist_cap) {
    const int N = A.N, m = p.m;
    const size_t ld = pad4(N);  // direction stride: every column 16-byte aligned
    if (m > kRedMaxK) return SENK_ERR_UNSUPPORTED;
    const bool prec = M.kind != 0 || M.learned;
    const int hcap = p.check_every ? p.outer * (m + 1) + 1 : p.outer + 1;
    size_t need = 2 * bytes_of<T>((size_t)(m + 1) * ld) + 4 * bytes_of<T>(N) + 3 * bytes_of<double>(m + 2) +
                  bytes_of<double>(hcap + 1) + bytes_of<int>(8);
    void *wsp;
    SENK_TRY(ctx_work(c, need, &wsp));
    Carver cv{(char *)wsp, 0, need};
    T *P = cv.take<T>((size_t)(m + 1) * ld), *AP = cv.take<T>((size_t)(m + 1) * ld);
    T *r = cv.take<T>(N), *Kr = cv.take<T>(N), *AKr = cv.take<T>(N), *tmp = cv.take<T>(N);
    double *dotAp = cv.take<double>(m + 2), *beta = cv.take<double>(m + 2);
    double *dhist = cv.take<double>(hcap + 1);
    int *state = cv.take<int>(8);
    Poller poll;
    SENK_TRY(poll.init(c, state));
    SolveTimer tm;
    SENK_TRY(tm.start(c));
    const double tol = p.nrm_b * p.eps;
    int converged = 0, breakdown = 0, last = 0, cycles = 0, nh = 0, hstate[4] = {0, 0, 0, 0};
    double restart_prev = 0;
    const int B = red_grid(N);
    for (int i = 0; i < p.outer; i++) {
        cycles++;
        int rg = 0;
        SENK_TRY(residual<T>(c, A, x, b, r, &rg));
        const int hidx = p.check_every ? i * (m + 1) : i;
        k_gcr_restart<<<1, kBlock, 0, c->stream>>>(red_slot(c, kSlotRes), rg, tol, p.nrm_b, state, dhist, hidx, hcap);
        SENK_LAUNCH_CHECK();
        SENK_HIP(hipMemcpyAsync(hstate, state, sizeof(int) * 4, hipMemcpyDeviceToHost, c->stream));
        if (M.learned && !M.learned->off) {
            // the restart residual is the true one: watch it for stagnation
            double rn = 0;
            SENK_HIP(hipMemcpyAsync(&rn, dhist + hidx, sizeof(double), hipMemcpyDeviceToHost, c->stream));
            SENK_HIP(hipStreamSynchronize(c->stream));
            if (i > 0) monitor_learned(M, rn, restart_prev);
            restart_prev = rn;
        }
        SENK_HIP(hipStreamSynchronize(c->stream));
        nh = hidx + 1;
        last = i * m;
        if (hstate[0]) {
            breakdown = hstate[2];
            converged = !breakdown;
            break;
        }
        // p_0 = M^-1 r (or r);  Ap_0 = A p_0
        if (prec) SENK_TRY(prec_apply<T>(c, M, r, nullptr, P, tmp, state));
        else SENK_TRY(launch_copy<T>(c, N, r, P));
        SENK_TRY(spmv<T>(c, A, P, AP, state));
        poll.reset();
        bool stop = false;
        for (int j = 0; j < m && !stop; j++) {
            T *pj = P + (size_t)j * ld, *Apj = AP + (size_t)j * ld;
            if (j == 0) {  // later steps get these partials from k_gcr_newdir
                k_gcr_alpha<T><<<B, kBlock, 0, c->stream>>>(N, Apj, r, red_slot(c, kSlotA), state);
                SENK_LAUNCH_CHECK();
            }
            k_gcr_xr<T><<<B, kBlock, 0, c->stream>>>(N, pj, Apj, red_slot(c, kSlotA), B, dotAp + j, x, r,
                                                     red_slot(c, kSlotC), state);
            SENK_LAUNCH_CHECK();
            if (p.check_every) {
                k_gcr_check<<<1, kBlock, 0, c->stream>>>(red_slot(c, kSlotC), B, tol, p.nrm_b, i * (m + 1) + 1 + j,
                                                         j, state, dhist, hcap);
                SENK_LAUNCH_CHECK();
            }
            const T *z = r;
            if (prec) {
                SENK_TRY(prec_apply<T>(c, M, r, nullptr, Kr, tmp, state));
                z = Kr;
            }
            SENK_TRY(spmv<T>(c, A, z, AKr, state));
            const int K = j + 1;
            if (K <= 8)
                k_gcr_betas<T, 8><<<B, kBlock, 0, c->stream>>>(N, K, AP, ld, AKr, red_slot(c, kSlotB), state);
            else if (K <= 16)
                k_gcr_betas<T, 16><<<B, kBlock, 0, c->stream>>>(N, K, AP, ld, AKr, red_slot(c, kSlotB), state);
            else if (K <= 24)
                k_gcr_betas<T, 24><<<B, kBlock, 0, c->stream>>>(N, K, AP, ld, AKr, red_slot(c, kSlotB), state);
            else if (K <= 32)
                k_gcr_betas<T, 32><<<B, kBlock, 0, c->stream>>>(N, K, AP, ld, AKr, red_slot(c, kSlotB), state);
            else
                k_gcr_betas<T, kRedMaxK><<<B, kBlock, 0, c->stream>>>(N, K, AP, ld, AKr, red_slot(c, kSlotB), state);
            SENK_LAUNCH_CHECK();
            k_gcr_beta_fin<<<K, kBlock, 0, c->stream>>>(red_slot(c, kSlotB), B, dotAp, beta, state);
            SENK_LAUNCH_CHECK();
            unsigned seq = 0;
            unsigned long long *mark = p.check_every ? poll.dev_mark(&seq) : nullptr;
            k_gcr_newdir<T><<<B, kBlock, 0, c->stream>>>(N, K, P, AP, ld, z, AKr, beta, r, red_slot(c, kSlotA), state,
                                                         mark, seq);
            SENK_LAUNCH_CHECK();
            if (p.check_every) SENK_TRY(poll.poll(kLag, &stop));
        }
        // a cycle that ran to its end did m steps
        last = (i + 1) * m;
        if (p.check_every) {
            SENK_HIP(hipMemcpyAsync(hstate, state, sizeof(int) * 4, hipMemcpyDeviceToHost, c->stream));
            SENK_HIP(hipStreamSynchronize(c->stream));
            nh = i * (m + 1) + 1 + hstate[1];
            if (hstate[0]) {
                breakdown = hstate[2];
                converged = !breakdown;
                last = i * m + hstate[1];
                break;
            }
        }
    }
    float ms = 0;
    SENK_TRY(tm.stop(c, &ms));
    std::vector<double> hh((size_t)hcap + 1, 0.0);
    SENK_HIP(hipMemcpy(hh.data(), dhist, sizeof(double) * hcap, hipMemcpyDeviceToHost));
    // history: one value per restart (reference), or restart value then one value per
    // step, cycle after cycle (check_every)
    const int out_n = std::min(nh, hcap);
    if (hist)
        for (int k = 0; k < std::min(out_n, hist_cap); k++) hist[k] = hh[k];
    res->iters = last;
    res->converged = converged;
    res->restarts = cycles;
    res->implicit_relres = out_n > 0 ? hh[out_n - 1] : 0;
    res->t_solve_ms = ms;
    res->status = breakdown ? SENK_ERR_BREAKDOWN : SENK_OK;
    if (p.verbose) {
        for (int k = 0; k < std::min(cycles, hcap); k++) {
            const int idx = p.check_every ? k * (m + 1) : k;
            std::printf("%d %e\n", k * m, hh[idx]);
        }
        std::fflush(stdout);
    }
    return SENK_OK;
}

// ================================================================= BiCGStab kernels
// Scalars in sc[]: [0] alpha, [1] omega, [2]/[3] rr (double-buffered by iteration
// parity so that a kernel never overwrites a value other workgroups still read).
namespace {
// partials of (AKp, rstr)                                       (senk_bicgstab.hpp:140)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_bicg_alpha(int N, const T *__restrict__ AKp, const T *__restrict__ rs,
                                                       double *Pout, const int *stop) {
    if (*stop) return;
    double a = 0;
    SENK_FOR4(i0, N) {
        T x[4], y[4];
        get4(AKp, i0, N, x);
        get4(rs, i0, N, y);
#pragma unroll
        for (int u = 0; u < 4; u++) a += (double)x[u] * (double)y[u];
    }
    const double v[1] = {a};
    store_partials<1>(v, Pout);
}
// alpha = rr / (AKp, rstr);  s = -alpha AKp + r                    (:140-141)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_bicg_s(int N, const T *__restrict__ AKp, const T *__restrict__ r,
                                                   const double *Pin, int B, const double *rr, double *alpha_out,
                                                   T *__restrict__ s, const int *stop) {
    if (*stop) return;
    const double alpha = *rr / sum_partials(Pin, B);
    if (blockIdx.x == 0 && threadIdx.x == 0) *alpha_out = alpha;
    const T a = (T)(-alpha);
    SENK_FOR4(i0, N) {
        T x[4], y[4];
        get4(AKp, i0, N, x);
        get4(r, i0, N, y);
#pragma unroll
        for (int u = 0; u < 4; u++) y[u] = a * x[u] + y[u];
        put4(s, i0, N, y);
    }
}
// partials of (AKs, s) and (AKs, AKs)                            (:145)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_bicg_omega(int N, const T *__restrict__ AKs, const T *__restrict__ s,
                                                       double *Pout, const int *stop) {
    if (*stop) return;
    double a0 = 0, a1 = 0;
    SENK_FOR4(i0, N) {
        T x[4], y[4];
        get4(AKs, i0, N, x);
        get4(s, i0, N, y);
#pragma unroll
        for (int u = 0; u < 4; u++) {
            a0 += (double)x[u] * (double)y[u];
            a1 += (double)x[u] * (double)x[u];
        }
    }
    const double v[2] = {a0, a1};
    store_partials<2>(v, Pout);
}
// omega = (AKs,s)/(AKs,AKs); x += alpha Kp; x += omega Ks; r = -omega AKs + s;
// partials of ||r||^2 and (r, rstr)                              (:145-158)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_bicg_update(int N, const T *__restrict__ Kp, const T *__restrict__ Ks,
                                                        const T *__restrict__ AKs, const T *__restrict__ s,
                                                        const T *__restrict__ rs, T *x, T *r, double *sc,
                                                        const double *Pin, int B, double *Pout, const int *stop) {
    if (*stop) return;
    const double omega = sum_partials(Pin, B) / sum_partials(Pin + kPartLd, B);
    if (blockIdx.x == 0 && threadIdx.x == 0) sc[1] = omega;
    const T a = (T)sc[0], w = (T)omega;
    const bool ks_is_s = Ks == s;  // the split form passes s for both: read it once
    double ss = 0, rr = 0;
    SENK_FOR4(i0, N) {
        T xv[4], t[4], rv[4];
        get4(x, i0, N, xv);
        get4(Kp, i0, N, t);
#pragma unroll
        for (int u = 0; u < 4; u++) xv[u] = xv[u] + a * t[u];
        get4(Ks, i0, N, t);
#pragma unroll
        for (int u = 0; u < 4; u++) xv[u] = xv[u] + w * t[u];
        put4(x, i0, N, xv);
        if (ks_is_s) {
#pragma unroll
            for (int u = 0; u < 4; u++) rv[u] = t[u];
        } else {
            get4(s, i0, N, rv);
        }
        get4(AKs, i0, N, t);
#pragma unroll
        for (int u = 0; u < 4; u++) rv[u] = (-w) * t[u] + rv[u];
        put4(r, i0, N, rv);
        get4(rs, i0, N, t);
#pragma unroll
        for (int u = 0; u < 4; u++) {
            ss += (double)rv[u] * (double)rv[u];
            rr += (double)rv[u] * (double)t[u];
        }
    }
    const double v[2] = {ss, rr};
    store_partials<2>(v, Pout);
}
// stop test on ||r||; beta = alpha/omega * rr_new/rr_old; p = beta(-omega AKp + p) + r
// (:149-161)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_bicg_p(int N, const T *__restrict__ AKp, const T *__restrict__ r,
                                                   double *sc, int parity, const double *Pin, int B, double tol,
                                                   double nrm_b, int iter, int *state, double *hist, int hcap,
                                                   T *p, unsigned long long *mark, unsigned seq,
                                                   const double *dtol = nullptr) {
    if (dtol) {  // tolerance and norm of the right-hand side formed on the device
        tol = dtol[0];
        nrm_b = dtol[1];
    }
    if (state[0]) {
        if (mark && blockIdx.x == 0 && threadIdx.x == 0) write_mark(mark, seq, 1);
        return;
    }
    const double nr = sqrt(sum_partials(Pin, B));
    const double rr_new = sum_partials(Pin + kPartLd, B);
    const bool fin = !isfinite(nr), conv = nr < tol;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (hist && iter < hcap) hist[iter] = nr / nrm_b;
        state[1] = iter + 1;
        if (fin) {
            state[2] = 1;
            state[0] = 1;
        } else if (conv) {
            state[0] = 1;
        } else {
            sc[2 + (parity ^ 1)] = rr_new;
        }
        if (mark) write_mark(mark, seq, fin || conv);
    }
    if (fin || conv) return;
    const double rr_old = sc[2 + parity];
    const double beta = sc[0] / sc[1] * rr_new / rr_old;
    const T w = (T)(-sc[1]), be = (T)beta;
    SENK_FOR4(i0, N) {
        T a[4], pv[4], rv[4];
        get4(AKp, i0, N, a);
        get4(p, i0, N, pv);
        get4(r, i0, N, rv);
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const T tmp = w * a[u] + pv[u];
            pv[u] = be * tmp + rv[u];
        }
        put4(p, i0, N, pv);
    }
}
__global__ void k_state_reset(int *state) {
    if (threadIdx.x == 0) {
        state[0] = 0;
        state[1] = 0;
        state[2] = 0;
    }
}
// rr = (r, rstr); with dtol: dtol = {eps ||r||, ||r||} (r = rstr); a zero r stops as
// converged, a non-finite one as a breakdown
__global__ __launch_bounds__(kBlock) void k_bicg_init(const double *P, int B, double *sc, int *state,
                                                      double eps = 0, double *dtol = nullptr) {
    const double rr = sum_partials(P, B);
    if (threadIdx.x != 0) return;
    sc[2] = rr;
    state[0] = 0;
    state[1] = 0;
    state[2] = 0;
    if (dtol) {
        dtol[0] = eps * sqrt(rr);
        dtol[1] = sqrt(rr);
        if (!isfinite(rr)) {
            state[0] = 1;
            state[2] = 1;
        } else if (!(rr > 0)) {
            state[0] = 1;
        }
    }
}
// Split form for a D-ILU factorization (Eisenstat's trick).  With M = L U (L unit lower,
// U upper with diagonal D) and A = L D + U - E, E = 2 D - diag(A):
//   L^-1 A U^-1 v = D t + L^-1 (v - E t),  t = U^-1 v,
// one U solve, one L solve and two vector passes, no product with A.
// w = v - E t
__global__ __launch_bounds__(kBlock) void k_dilu_pre(int N, const float *__restrict__ v, const float *__restrict__ e,
                                                     const float *__restrict__ t, float *w, const int *stop) {
    if (*stop) return;
    SENK_FOR4(i0, N) {
        float a[4], b[4], ev[4];
        get4(v, i0, N, a);
        get4(e, i0, N, ev);
        get4(t, i0, N, b);
#pragma unroll
        for (int u = 0; u < 4; u++) a[u] = a[u] - ev[u] * b[u];
        put4(w, i0, N, a);
    }
}
// out = D t + y, with the partial sums of the BiCGStab step's dot products of out:
// K = 1 (out, rs) (as k_bicg_alpha), K = 2 (out, s) and (out, out) (as k_bicg_omega); the
// same grid and loop give the same partials as those kernels
template <int K>
__global__ __launch_bounds__(kBlock) void k_dilu_post(int N, const float *__restrict__ d, const float *__restrict__ t,
                                                      const float *__restrict__ y, float *out,
                                                      const float *__restrict__ other, double *Pout,
                                                      const int *stop) {
    if (*stop) return;
    double a0 = 0, a1 = 0;
    SENK_FOR4(i0, N) {
        float dv[4], tv[4], yv[4], ov[4];
        get4(d, i0, N, dv);
        get4(t, i0, N, tv);
        get4(y, i0, N, yv);
#pragma unroll
        for (int u = 0; u < 4; u++) yv[u] = dv[u] * tv[u] + yv[u];
        put4(out, i0, N, yv);
        get4(other, i0, N, ov);
#pragma unroll
        for (int u = 0; u < 4; u++) {
            a0 += (double)yv[u] * (double)ov[u];
            if (K == 2) a1 += (double)yv[u] * (double)yv[u];
        }
    }
    if (K == 1) {
        const double v[1] = {a0};
        store_partials<1>(v, Pout);
    } else {
        const double v[2] = {a0, a1};
        store_partials<2>(v, Pout);
    }
}
}  // namespace

namespace {
int read_scalar(Ctx *c, const double *d, double *h);
}

// v = L^-1 A U^-1 in with the layouts (t, w scratch; w may be v; the U sweep writes
// w = in - E t itself where it runs as one launch), and the partials of
// (v, other) (K = 1) or of (v, other), (v, v) (K = 2) over red_grid(N) workgroups
template <int K>
static int dilu_product(Ctx *c, const PrecView &M, const float *in, float *v, float *t, float *w,
                        const float *other, double *Pout, const int *stop) {
    const int N = M.N;
    bool fused = false;
    SENK_TRY(launch_tri_single<float>(c, M.uts, *M.ut, *M.lt, in, t, stop, M.dilu_e, w, &fused));
    if (!fused) {
        k_dilu_pre<<<ew_grid(N), kBlock, 0, c->stream>>>(N, in, M.dilu_e, t, w, stop);
        SENK_LAUNCH_CHECK();
    }
    SENK_TRY(launch_tri_single<float>(c, M.lts, *M.lt, *M.ut, w, w, stop));
    k_dilu_post<K><<<red_grid(N), kBlock, 0, c->stream>>>(N, M.dilu_d, t, w, v, other, Pout, stop);
    SENK_LAUNCH_CHECK();
    return SENK_OK;
}

// BiCGStab on L^-1 A U^-1 y = L^-1 b from y = 0 (the BiCGStab steps of
// senk_bicgstab.hpp:129-161 with the operator above and no further preconditioner),
// x = U^-1 y.  x must be zero on entry; the residual it stops on is L^-1 (b - A x),
// relative to ||L^-1 b||.  Every triangle alternates L, U as the layouts require:
// L (start), (U, L) per product, U (end).
static int bicgstab_split(Ctx *c, const MatView &A, const PrecView &M, const float *b, float *x,
                          const SolveParams &p, senk_result *res, double *hist, int hist_cap) {
    const int N = A.N;
    const int hcap = p.max_iter + 1;
    size_t need = 10 * bytes_of<float>(N) + bytes_of<double>(hcap + 1) + bytes_of<int>(8) + bytes_of<double>(16);
    void *wsp;
    SENK_TRY(ctx_work(c, need, &wsp));
    Carver cv{(char *)wsp, 0, need};
    float *r = cv.take<float>(N), *rs = cv.take<float>(N), *pp = cv.take<float>(N), *v = cv.take<float>(N);
    float *s = cv.take<float>(N), *tv = cv.take<float>(N), *y = cv.take<float>(N), *t = cv.take<float>(N);
    float *w = cv.take<float>(N);
    double *dhist = cv.take<double>(hcap + 1);
    int *state = cv.take<int>(8);
    double *sc = cv.take<double>(16);
    Poller poll;
    SENK_TRY(poll.init(c, state));
    SolveTimer tm;
    SENK_TRY(tm.start(c));
    const int B = red_grid(N);
    // the state words come from the shared arena: clear the stop flag before the sweeps
    k_state_reset<<<1, 64, 0, c->stream>>>(state);
    SENK_LAUNCH_CHECK();
    // r = L^-1 b; rstr = r; p = r; y = 0; rr = (r, rstr)
    SENK_TRY(launch_tri_single<float>(c, M.lts, *M.lt, *M.ut, b, r, state));
    SENK_TRY(launch_copy<float>(c, N, r, rs));
    SENK_TRY(launch_copy<float>(c, N, r, pp));
    SENK_HIP(hipMemsetAsync(y, 0, sizeof(float) * N, c->stream));
    k_bicg_alpha<float><<<B, kBlock, 0, c->stream>>>(N, r, rs, red_slot(c, kSlotA), state);
    SENK_LAUNCH_CHECK();
    double *dtol = sc + 8;  // {tolerance, ||L^-1 b||}, formed on the device (no host round trip)
    k_bicg_init<<<1, kBlock, 0, c->stream>>>(red_slot(c, kSlotA), B, sc, state, p.eps, dtol);
    SENK_LAUNCH_CHECK();
    poll.reset();
    bool stop = false;
    for (int i = 0; i < p.max_iter && !stop; i++) {
        const int par = i & 1;
        SENK_TRY(dilu_product<1>(c, M, pp, v, t, w, rs, red_slot(c, kSlotA), state));
        k_bicg_s<float><<<B, kBlock, 0, c->stream>>>(N, v, r, red_slot(c, kSlotA), B, sc + 2 + par, sc + 0, s, state);
        SENK_LAUNCH_CHECK();
        SENK_TRY(dilu_product<2>(c, M, s, tv, t, w, s, red_slot(c, kSlotB), state));
        k_bicg_update<float><<<B, kBlock, 0, c->stream>>>(N, pp, s, tv, s, rs, y, r, sc, red_slot(c, kSlotB), B,
                                                          red_slot(c, kSlotC), state);
        SENK_LAUNCH_CHECK();
        unsigned seq = 0;
        unsigned long long *mark = poll.dev_mark(&seq);
        k_bicg_p<float><<<B, kBlock, 0, c->stream>>>(N, v, r, sc, par, red_slot(c, kSlotC), B, 0.0, 1.0, i, state,
                                                     dhist, hcap, pp, mark, seq, dtol);
        SENK_LAUNCH_CHECK();
        SENK_TRY(poll.poll(kLag, &stop));
    }
    // x = U^-1 y (after the stop flag is set: no flag for this sweep)
    SENK_TRY(launch_tri_single<float>(c, M.uts, *M.ut, *M.lt, y, x, nullptr));
    int hstate[4];
    SENK_HIP(hipMemcpyAsync(hstate, state, sizeof(int) * 4, hipMemcpyDeviceToHost, c->stream));
    float ms = 0;
    SENK_TRY(tm.stop(c, &ms));
    const int converged = hstate[0] && !hstate[2];
    const int iters = hstate[0] ? hstate[1] : p.max_iter;
    std::vector<double> hh((size_t)std::max(iters, 1), 0.0);
    if (iters > 0)
        SENK_HIP(hipMemcpy(hh.data(), dhist, sizeof(double) * std::min(iters, hcap), hipMemcpyDeviceToHost));
    if (hist)
        for (int k = 0; k < std::min(iters, hist_cap); k++) hist[k] = hh[k];
    res->iters = iters;
    res->converged = converged;
    res->restarts = 0;
    res->implicit_relres = iters > 0 ? hh[iters - 1] : 0;
    res->t_solve_ms = ms;
    res->status = hstate[2] ? SENK_ERR_BREAKDOWN : SENK_OK;
    return SENK_OK;
}

template <typename T>
static int bicgstab(Ctx *c, const MatView &A, const PrecView &M, const T *b, T *x, const SolveParams &p,
                    senk_result *res, double *hist, int hist_cap) {
    if constexpr (std::is_same<T, float>::value) {
        if (p.split && M.dilu_d && M.dilu_e && M.lt && M.ut && !M.f16)
            return bicgstab_split(c, A, M, b, x, p, res, hist, hist_cap);
    }
    const int N = A.N;
    const int hcap = p.max_iter + 1;
    size_t need = 10 * bytes_of<T>(N) + bytes_of<double>(hcap + 1) + bytes_of<int>(8) + bytes_of<double>(16);
    void *wsp;
    SENK_TRY(ctx_work(c, need, &wsp));
    Carver cv{(char *)wsp, 0, need};
    T *r = cv.take<T>(N), *rs = cv.take<T>(N), *pp = cv.take<T>(N), *Kp = cv.take<T>(N), *AKp = cv.take<T>(N);
    T *s = cv.take<T>(N), *Ks = cv.take<T>(N), *AKs = cv.take<T>(N), *tmp = cv.take<T>(N);
    double *dhist = cv.take<double>(hcap + 1);
    int *state = cv.take<int>(8);
    double *sc = cv.take<double>(16);
    Poller poll;
    SENK_TRY(poll.init(c, state));
    SolveTimer tm;
    SENK_TRY(tm.start(c));
    const double tol = p.eps * p.nrm_b;
    const int B = red_grid(N);
    // r = b - A x; rstr = r; p = r; rr = (r, rstr)
    int rg = 0;
    SENK_TRY(residual<T>(c, A, x, b, r, &rg));
    SENK_TRY(launch_copy<T>(c, N, r, rs));
    SENK_TRY(launch_copy<T>(c, N, r, pp));
    // the state words come from the shared arena: clear the stop flag before the first
    // flag-checking kernel
    k_state_reset<<<1, 64, 0, c->stream>>>(state);
    SENK_LAUNCH_CHECK();
    k_bicg_alpha<T><<<B, kBlock, 0, c->stream>>>(N, r, rs, red_slot(c, kSlotA), state);
    SENK_LAUNCH_CHECK();
    k_bicg_init<<<1, kBlock, 0, c->stream>>>(red_slot(c, kSlotA), B, sc, state);
    SENK_LAUNCH_CHECK();
    poll.reset();
    bool stop = false;
    for (int i = 0; i < p.max_iter && !stop; i++) {
        const int par = i & 1;
        SENK_TRY(prec_apply<T>(c, M, pp, nullptr, Kp, tmp, state));
        SENK_TRY(spmv<T>(c, A, Kp, AKp, state));
        k_bicg_alpha<T><<<B, kBlock, 0, c->stream>>>(N, AKp, rs, red_slot(c, kSlotA), state);
        SENK_LAUNCH_CHECK();
        k_bicg_s<T><<<B, kBlock, 0, c->stream>>>(N, AKp, r, red_slot(c, kSlotA), B, sc + 2 + par, sc + 0, s, state);
        SENK_LAUNCH_CHECK();
        SENK_TRY(prec_apply<T>(c, M, s, nullptr, Ks, tmp, state));
        SENK_TRY(spmv<T>(c, A, Ks, AKs, state));
        k_bicg_omega<T><<<B, kBlock, 0, c->stream>>>(N, AKs, s, red_slot(c, kSlotB), state);
        SENK_LAUNCH_CHECK();
        k_bicg_update<T><<<B, kBlock, 0, c->stream>>>(N, Kp, Ks, AKs, s, rs, x, r, sc, red_slot(c, kSlotB), B,
                                                      red_slot(c, kSlotC), state);
        SENK_LAUNCH_CHECK();
        unsigned seq = 0;
        unsigned long long *mark = poll.dev_mark(&seq);
        k_bicg_p<T><<<B, kBlock, 0, c->stream>>>(N, AKp, r, sc, par, red_slot(c, kSlotC), B, tol, p.nrm_b, i, state,
                                                 dhist, hcap, pp, mark, seq);
        SENK_LAUNCH_CHECK();
        SENK_TRY(poll.poll(kLag, &stop));
    }
    int hstate[4];
    SENK_HIP(hipMemcpyAsync(hstate, state, sizeof(int) * 4, hipMemcpyDeviceToHost, c->stream));
    float ms = 0;
    SENK_TRY(tm.stop(c, &ms));
    const int converged = hstate[0] && !hstate[2];
    const int iters = hstate[0] ? hstate[1] : p.max_iter;
    std::vector<double> hh((size_t)std::max(iters, 1), 0.0);
    if (iters > 0)
        SENK_HIP(hipMemcpy(hh.data(), dhist, sizeof(double) * std::min(iters, hcap), hipMemcpyDeviceToHost));
    if (hist)
        for (int k = 0; k < std::min(iters, hist_cap); k++) hist[k] = hh[k];
    res->iters = iters;
    res->converged = converged;
    res->restarts = 0;
    res->implicit_relres = iters > 0 ? hh[iters - 1] : 0;
    res->t_solve_ms = ms;
    res->status = hstate[2] ? SENK_ERR_BREAKDOWN : SENK_OK;
    if (p.verbose) {
        for (int k = 0; k < iters; k++) std::printf("%d %e\n", k + 1, hh[k]);
        if (converged) {
            std::printf("# iter %d\n", iters);
            std::printf("# res %e\n", res->implicit_relres);
        } else {
            std::printf("# iter %d (max)\n", iters);
            std::printf("# res %e\n", res->implicit_relres);
        }
        std::fflush(stdout);
    }
    return SENK_OK;
}

// ================================================================= CG
// Preconditioned conjugate gradients.  The reference announces CG (README_ja.md:44,
// docs/index.html:174-175) but has none; these are the textbook recurrences with the
// reference BiCGStab's stop test ||r|| < eps ||b|| and count.  M must be symmetric
// positive definite (IC(0)/IC(p): L D L^T).  Scalars in sc[]: [2]/[3] (r, z),
// double-buffered by iteration parity.
namespace {
// alpha = (r, z) / (p, q);  x += alpha p;  r -= alpha q;  partials of ||r||^2
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_update(int N, const T *__restrict__ p, const T *__restrict__ q,
                                                      T *x, T *r, const double *rz, const double *Pin, int B,
                                                      double *Pout, const int *stop) {
    if (*stop) return;
    const double alpha = *rz / sum_partials(Pin, B);
    const T a = (T)alpha;
    double rr = 0;
    SENK_FOR4(i0, N) {
        T pv[4], xv[4], qv[4], rv[4];
        get4(p, i0, N, pv);
        get4(x, i0, N, xv);
#pragma unroll
        for (int u = 0; u < 4; u++) xv[u] = a * pv[u] + xv[u];
        put4(x, i0, N, xv);
        get4(q, i0, N, qv);
        get4(r, i0, N, rv);
#pragma unroll
        for (int u = 0; u < 4; u++) {
            rv[u] = (-a) * qv[u] + rv[u];
            rr += (double)rv[u] * (double)rv[u];
        }
        put4(r, i0, N, rv);
    }
    const double v[1] = {rr};
    store_partials<1>(v, Pout);
}
// sc[2] = (r, z); a zero (r, z) (r = 0: x already solves the system) stops as converged
// with x unchanged, a negative or non-finite one (M not positive definite) as a breakdown
__global__ __launch_bounds__(kBlock) void k_cg_init(const double *P, int B, double *sc, int *state) {
    const double rz = sum_partials(P, B);
    if (threadIdx.x != 0) return;
    sc[2] = rz;
    state[0] = 0;
    state[1] = 0;
    state[2] = 0;
    if (!isfinite(rz) || rz < 0) {
        state[0] = 1;
        state[2] = 1;
    } else if (rz == 0) {
        state[0] = 1;
    }
}
// stop test on ||r||; beta = (r, z)_new / (r, z)_old;  p = z + beta p
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_p(int N, const T *__restrict__ z, T *p, double *sc, int parity,
                                                 const double *Prr, const double *Prz, int B, double tol,
                                                 double nrm_b, int iter, int *state, double *hist, int hcap,
                                                 unsigned long long *mark, unsigned seq) {
    if (state[0]) {
        if (mark && blockIdx.x == 0 && threadIdx.x == 0) write_mark(mark, seq, 1);
        return;
    }
    const double nr = sqrt(sum_partials(Prr, B));
    const double rz_new = sum_partials(Prz, B);
    const bool fin = !isfinite(nr) || !isfinite(rz_new), conv = nr < tol;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (hist && iter < hcap) hist[iter] = nr / nrm_b;
        state[1] = iter + 1;
        if (fin) {
            state[2] = 1;
            state[0] = 1;
        } else if (conv) {
            state[0] = 1;
        } else {
            sc[2 + (parity ^ 1)] = rz_new;
        }
        if (mark) write_mark(mark, seq, fin || conv);
    }
    if (fin || conv) return;
    const T be = (T)(rz_new / sc[2 + parity]);
    SENK_FOR4(i0, N) {
        T zv[4], pv[4];
        get4(z, i0, N, zv);
        get4(p, i0, N, pv);
#pragma unroll
        for (int u = 0; u < 4; u++) pv[u] = zv[u] + be * pv[u];
        put4(p, i0, N, pv);
    }
}
}  // namespace

// Split form of CG for a D-ILU factorization of a symmetric A (IC(0), or ILU(0) on a
// triangle-free graph: U = D L^T to rounding).  PCG with M = L D L^T is CG on
// B = L^-1 A L^-T (y = L^T x, c = L^-1 b) preconditioned by D^-1.  Carrying u = D p and
// yd = D y instead of p and y: B p = L^-1 A U^-1 u (the split product: one U and one L
// solve, no product with A), D z = r, so u <- r + beta u, yd <- yd + alpha u, and at the
// end x = U^-1 yd.  The residual it stops on is L^-1 (b - A x), relative to ||L^-1 b||.
// Per step: the U sweep (taking the previous step's yd, r, u update as it reads u, and
// with w = u - E t in its epilogue), the L sweep (with q = D t + L^-1 w in its epilogue),
// one pass for all seven dot products the step needs and a one-workgroup kernel for the
// scalars; the new (r, r/d) and (r, r) come from the expansion of r - alpha q.  No vector
// pass of its own for the update: 128^3 nested solve 20.8 ms against 21.1 ms with one
// (a pass that also summed the partials in every workgroup), the U sweep taking 27.6
// instead of 22.4 us.
namespace {
// partials of (r, r) and (r, r/d)
__global__ __launch_bounds__(kBlock) void k_cgs_start(int N, const float *__restrict__ r,
                                                      const float *__restrict__ d, double *Pout, const int *stop) {
    if (*stop) return;
    double a0 = 0, a1 = 0;
    SENK_FOR4(i0, N) {
        float rv[4], dv[4];
        get4(r, i0, N, rv);
        get4_fill(d, i0, N, dv, 1.0f);
#pragma unroll
        for (int u = 0; u < 4; u++) {
            a0 += (double)rv[u] * (double)rv[u];
            a1 += (double)rv[u] * (double)rv[u] / (double)dv[u];
        }
    }
    const double v[2] = {a0, a1};
    store_partials<2>(v, Pout);
}
// sc[2] = rho = (r, r/d); dtol = {eps ||r||, ||r||}; a zero r stops as converged, a
// non-finite one or a negative rho (a pivot d < 0: M not positive definite) as a breakdown
__global__ __launch_bounds__(kBlock) void k_cgs_init(const double *P, int B, double *sc, int *state, double eps,
                                                     double *dtol) {
    const double rr = sum_partials(P, B), rho = sum_partials(P + kPartLd, B);
    if (threadIdx.x != 0) return;
    sc[2] = rho;
    state[0] = 0;
    state[1] = 0;
    state[2] = 0;
    dtol[0] = eps * sqrt(rr);
    dtol[1] = sqrt(rr);
    if (!isfinite(rr) || !isfinite(rho) || rho < 0) {
        state[0] = 1;
        state[2] = 1;
    } else if (!(rr > 0)) {
        state[0] = 1;
    }
}
// q = D t + q (the product when the L sweep did not form it)
__global__ __launch_bounds__(kBlock) void k_cgs_q(int N, const float *__restrict__ d, const float *__restrict__ t,
                                                  float *q, const int *stop) {
    if (*stop) return;
    SENK_FOR4(i0, N) {
        float dv[4], tv[4], qv[4];
        get4(d, i0, N, dv);
        get4(t, i0, N, tv);
        get4(q, i0, N, qv);
#pragma unroll
        for (int k = 0; k < 4; k++) qv[k] = dv[k] * tv[k] + qv[k];
        put4(q, i0, N, qv);
    }
}
// One reading pass for the whole step: the partials of (q, u/d), (r, r/d), (q, r/d),
// (q, q/d), (r, r), (q, r), (q, q).
constexpr int kCgsDots = 7;
constexpr int kCgsMaxB = 2 * kBlock;  // workgroups of k_cgs_dots: two partials per thread in k_cgs_scalars
__global__ __launch_bounds__(kBlock) void k_cgs_dots(int N, const float *__restrict__ q,
                                                     const float *__restrict__ u, const float *__restrict__ r,
                                                     const float *__restrict__ d, double *P, const int *state) {
    if (state[0]) return;
    double a[kCgsDots] = {};
    SENK_FOR4(i0, N) {
        float qv[4], uv[4], rv[4], dv[4];
        get4(q, i0, N, qv);
        get4(u, i0, N, uv);
        get4(r, i0, N, rv);
        get4_fill(d, i0, N, dv, 1.0f);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const double qk = qv[k], uk = uv[k], rk = rv[k], id = 1.0 / (double)dv[k];
            const double qd = qk * id, rd = rk * id;
            a[0] += qd * uk;
            a[1] += rd * rk;
            a[2] += qd * rk;
            a[3] += qd * qk;
            a[4] += rk * rk;
            a[5] += qk * rk;
            a[6] += qk * qk;
        }
    }
    store_partials<kCgsDots>(a, P);
}
// The step's scalars from those sums (one workgroup): alpha = (r, r/d) / (q, u/d); the
// new residual's (r, r/d) and (r, r) by expansion, r' = r - alpha q giving (r,r/d) -
// 2 alpha (q,r/d) + alpha^2 (q,q/d); the stop test on ||r'|| (tolerance formed on the
// device); beta = (r', r'/d) / (r, r/d).  ab = {alpha, beta} for the vector step
// yd += alpha u, r = r', u = r' + beta u, which the next U sweep does as it reads u
// (SweepPre) and k_cgs_last does for the step that ends the loop.  The partials are read
// before the stop flag, all at once.  (Formed instead by the last of k_cgs_dots'
// workgroups, counted with an atomic: 128^3 nested solve 21.8 ms with the partials
// passed through the device-coherent level, 26.3 ms with device-scope fences, which
// write back and invalidate an XCD's L2, against 20.8 ms for this kernel.)
__global__ __launch_bounds__(kBlock) void k_cgs_scalars(const double *__restrict__ P, int B,
                                                        const double *__restrict__ dtol, int iter, int *state,
                                                        float *ab, double *hist, int hcap,
                                                        unsigned long long *mark, unsigned seq) {
    double pv[2][kCgsDots];
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const int b = threadIdx.x + j * kBlock;
#pragma unroll
        for (int k = 0; k < kCgsDots; k++) pv[j][k] = b < B ? P[(size_t)k * kPartLd + b] : 0.0;
    }
    if (state[0]) {  // stopped at an earlier step
        if (mark && threadIdx.x == 0) write_mark(mark, seq, 1);
        return;
    }
    // the sums in slot order, as one pass over the partials forms them
    __shared__ double sh[kCgsDots][kWavesPerBlock];
    const int wv = threadIdx.x >> 6;
    double s[kCgsDots];
#pragma unroll
    for (int k = 0; k < kCgsDots; k++) {
        const double v = wave_sum(pv[0][k] + pv[1][k]);
        if ((threadIdx.x & 63) == 0) sh[k][wv] = v;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
#pragma unroll
    for (int k = 0; k < kCgsDots; k++) {
        double t = 0;
#pragma unroll
        for (int i = 0; i < kWavesPerBlock; i++) t += sh[k][i];
        s[k] = t;
    }
    const double sigma = s[0], rho = s[1];
    const double alpha = rho / sigma;
    const double rho_new = rho - 2 * alpha * s[2] + alpha * alpha * s[3];
    const double rr_new = s[4] - 2 * alpha * s[5] + alpha * alpha * s[6];
    const double nr = sqrt(rr_new > 0 ? rr_new : 0.0);
    const bool conv = nr < dtol[0];
    // breakdown: a non-finite sum, (q, u/d) <= 0 (the split operator is not positive
    // definite: a pivot d < 0), or (r', r'/d) < 0 short of convergence
    const bool fin = !isfinite(sigma) || !isfinite(rho_new) || !isfinite(rr_new) || !(sigma > 0) ||
                     (rho_new < 0 && !conv);
    if (hist && iter < hcap) hist[iter] = nr / dtol[1];
    state[1] = iter + 1;
    if (fin) {
        state[2] = 1;
        state[0] = 1;
    } else if (conv) {
        state[0] = 1;
    }
    ab[0] = (float)alpha;
    ab[1] = (float)(rho_new / rho);
    if (mark) write_mark(mark, seq, fin || conv);
}
// The vector step of the loop's last step (stopped by convergence or the step limit; a
// breakdown takes no step): yd += alpha u.
__global__ __launch_bounds__(kBlock) void k_cgs_last(int N, const float *__restrict__ u, float *yd,
                                                     const float *__restrict__ ab, const int *__restrict__ state) {
    if (state[2] || state[1] == 0) return;
    const float a = ab[0];
    SENK_FOR4(i0, N) {
        float uv[4], yv[4];
        get4(u, i0, N, uv);
        get4(yd, i0, N, yv);
#pragma unroll
        for (int k = 0; k < 4; k++) yv[k] = a * uv[k] + yv[k];
        put4(yd, i0, N, yv);
    }
}
}  // namespace

// x = U^-1 yd; x must be zero on entry.  Triangles alternate L, U as the layouts require:
// L (start), (U, L) per product, U (end).
static int cg_split(Ctx *c, const MatView &A, const PrecView &M, const float *b, float *x, const SolveParams &p,
                    senk_result *res, double *hist, int hist_cap) {
    const int N = A.N;
    const int hcap = p.max_iter + 1;
    size_t need = 6 * bytes_of<float>(N) + bytes_of<double>(hcap + 1) + bytes_of<int>(8) + bytes_of<double>(16);
    void *wsp;
    SENK_TRY(ctx_work(c, need, &wsp));
    Carver cv{(char *)wsp, 0, need};
    float *r = cv.take<float>(N), *u = cv.take<float>(N), *q = cv.take<float>(N), *t = cv.take<float>(N);
    float *w = cv.take<float>(N), *yd = cv.take<float>(N);
    double *dhist = cv.take<double>(hcap + 1);
    int *state = cv.take<int>(8);
    double *sc = cv.take<double>(16);
    double *dtol = sc + 8;  // {tolerance, ||L^-1 b||}, formed on the device
    float *ab = reinterpret_cast<float *>(sc + 12);  // the step's {alpha, beta}
    Poller poll;
    SENK_TRY(poll.init(c, state));
    SolveTimer tm;
    SENK_TRY(tm.start(c));
    // at most 512 workgroups per reducing pass (128^3 nested solve: 31.0 -> 30.8 ms, the
    // same iterations; 256 was slower)
    const int B = std::min(red_grid(N), kCgsMaxB);
    k_state_reset<<<1, 64, 0, c->stream>>>(state);
    SENK_LAUNCH_CHECK();
    // r = L^-1 b; u = r; yd = 0; rho = (r, r/d)
    SENK_TRY(launch_tri_single<float>(c, M.lts, *M.lt, *M.ut, b, r, state));
    SENK_TRY(launch_copy<float>(c, N, r, u));
    SENK_HIP(hipMemsetAsync(yd, 0, sizeof(float) * N, c->stream));
    k_cgs_start<<<B, kBlock, 0, c->stream>>>(N, r, M.dilu_d, red_slot(c, kSlotB), state);
    SENK_LAUNCH_CHECK();
    k_cgs_init<<<1, kBlock, 0, c->stream>>>(red_slot(c, kSlotB), B, sc, state, p.eps, dtol);
    SENK_LAUNCH_CHECK();
    poll.reset();
    // Where both triangles are solved by single-launch sweeps, t = U^-1 u is read by the L
    // sweep's epilogue from the U sweep's own vector (before the L sweep returns it to the
    // sentinel) and is not copied out
    float *umid = tri_single_sweep_vector<float>(*M.lt, *M.ut) ? tri_single_sweep_vector<float>(*M.ut, *M.lt)
                                                                : nullptr;
    bool stop = false;
    for (int i = 0; i < p.max_iter && !stop; i++) {
        // q = L^-1 A U^-1 u  (t = U^-1 u; w = u - E t; q = D t + L^-1 w), each vector pass
        // formed in a sweep's epilogue where the launch mode allows; from the second step on
        // the U solve first takes the previous step's yd, r, u update
        bool fused = false;
        SweepPre<float> pre;
        if (i > 0) pre = SweepPre<float>{u, r, q, yd, ab};
        SENK_TRY(launch_tri_single<float>(c, M.uts, *M.ut, *M.lt, u, umid ? nullptr : t, state, M.dilu_e, w, &fused,
                                          pre));
        if (!fused) {
            k_dilu_pre<<<ew_grid(N), kBlock, 0, c->stream>>>(N, u, M.dilu_e, t, w, state);
            SENK_LAUNCH_CHECK();
        }
        SENK_TRY(launch_tri_single<float>(c, M.lts, *M.lt, *M.ut, w, q, state, M.dilu_d, umid ? umid : t, &fused));
        if (!fused) {
            k_cgs_q<<<ew_grid(N), kBlock, 0, c->stream>>>(N, M.dilu_d, t, q, state);
            SENK_LAUNCH_CHECK();
        }
        k_cgs_dots<<<B, kBlock, 0, c->stream>>>(N, q, u, r, M.dilu_d, red_slot(c, kSlotA), state);
        SENK_LAUNCH_CHECK();
        unsigned seq = 0;
        unsigned long long *mark = poll.dev_mark(&seq);
        k_cgs_scalars<<<1, kBlock, 0, c->stream>>>(red_slot(c, kSlotA), B, dtol, i, state, ab, dhist, hcap, mark,
                                                   seq);
        SENK_LAUNCH_CHECK();
        SENK_TRY(poll.poll(kLag, &stop));
    }
    k_cgs_last<<<ew_grid(N), kBlock, 0, c->stream>>>(N, u, yd, ab, state);
    SENK_LAUNCH_CHECK();
    // x = U^-1 yd (after the stop flag is set: no flag for this sweep)
    SENK_TRY(launch_tri_single<float>(c, M.uts, *M.ut, *M.lt, yd, x, nullptr));
    int hstate[4];
    SENK_HIP(hipMemcpyAsync(hstate, state, sizeof(int) * 4, hipMemcpyDeviceToHost, c->stream));
    float ms = 0;
    SENK_TRY(tm.stop(c, &ms));
    const int converged = hstate[0] && !hstate[2];
    const int iters = hstate[0] ? hstate[1] : p.max_iter;
    std::vector<double> hh((size_t)std::max(iters, 1), 0.0);
    if (iters > 0)
        SENK_HIP(hipMemcpy(hh.data(), dhist, sizeof(double) * std::min(iters, hcap), hipMemcpyDeviceToHost));
    if (hist)
        for (int k = 0; k < std::min(iters, hist_cap); k++) hist[k] = hh[k];
    res->iters = iters;
    res->converged = converged;
    res->restarts = 0;
    res->implicit_relres = iters > 0 ? hh[iters - 1] : 0;
    res->t_solve_ms = ms;
    res->status = hstate[2] ? SENK_ERR_BREAKDOWN : SENK_OK;
    return SENK_OK;
}

template <typename T>
static int cg(Ctx *c, const MatView &A, const PrecView &M, const T *b, T *x, const SolveParams &p,
              senk_result *res, double *hist, int hist_cap) {
    if (M.learned) return SENK_ERR_ARG;  // CG needs a fixed, symmetric preconditioner
    if constexpr (std::is_same<T, float>::value) {
        if (p.split && M.dilu_d && M.dilu_e && M.lt && M.ut && !M.f16 && M.symmetric)
            return cg_split(c, A, M, b, x, p, res, hist, hist_cap);
    }
    const int N = A.N;
    const int hcap = p.max_iter + 1;
    size_t need = 5 * bytes_of<T>(N) + bytes_of<double>(hcap + 1) + bytes_of<int>(8) + bytes_of<double>(16);
    void *wsp;
    SENK_TRY(ctx_work(c, need, &wsp));
    Carver cv{(char *)wsp, 0, need};
    T *r = cv.take<T>(N), *z = cv.take<T>(N), *pp = cv.take<T>(N), *q = cv.take<T>(N), *tmp = cv.take<T>(N);
    double *dhist = cv.take<double>(hcap + 1);
    int *state = cv.take<int>(8);
    double *sc = cv.take<double>(16);
    Poller poll;
    SENK_TRY(poll.init(c, state));
    SolveTimer tm;
    SENK_TRY(tm.start(c));
    const double tol = p.eps * p.nrm_b;
    const int B = red_grid(N);
    // r = b - A x; z = M r; p = z; (r, z)
    int rg = 0;
    SENK_TRY(residual<T>(c, A, x, b, r, &rg));
    // the state words come from the shared arena: clear the stop flag first
    k_state_reset<<<1, 64, 0, c->stream>>>(state);
    SENK_LAUNCH_CHECK();
    SENK_TRY(prec_apply<T>(c, M, r, nullptr, z, tmp, state));
    SENK_TRY(launch_copy<T>(c, N, z, pp));
    k_bicg_alpha<T><<<B, kBlock, 0, c->stream>>>(N, r, z, red_slot(c, kSlotC), state);
    SENK_LAUNCH_CHECK();
    k_cg_init<<<1, kBlock, 0, c->stream>>>(red_slot(c, kSlotC), B, sc, state);
    SENK_LAUNCH_CHECK();
    poll.reset();
    bool stop = false;
    for (int i = 0; i < p.max_iter && !stop; i++) {
        const int par = i & 1;
        SENK_TRY(spmv<T>(c, A, pp, q, state));
        k_bicg_alpha<T><<<B, kBlock, 0, c->stream>>>(N, pp, q, red_slot(c, kSlotA), state);  // (p, q)
        SENK_LAUNCH_CHECK();
        k_cg_update<T><<<B, kBlock, 0, c->stream>>>(N, pp, q, x, r, sc + 2 + par, red_slot(c, kSlotA), B,
                                                    red_slot(c, kSlotB), state);
        SENK_LAUNCH_CHECK();
        SENK_TRY(prec_apply<T>(c, M, r, nullptr, z, tmp, state));
        k_bicg_alpha<T><<<B, kBlock, 0, c->stream>>>(N, r, z, red_slot(c, kSlotC), state);  // (r, z)
        SENK_LAUNCH_CHECK();
        unsigned seq = 0;
        unsigned long long *mark = poll.dev_mark(&seq);
        k_cg_p<T><<<B, kBlock, 0, c->stream>>>(N, z, pp, sc, par, red_slot(c, kSlotB), red_slot(c, kSlotC), B, tol,
                                               p.nrm_b, i, state, dhist, hcap, mark, seq);
        SENK_LAUNCH_CHECK();
        SENK_TRY(poll.poll(kLag, &stop));
    }
    int hstate[4];
    SENK_HIP(hipMemcpyAsync(hstate, state, sizeof(int) * 4, hipMemcpyDeviceToHost, c->stream));
    float ms = 0;
    SENK_TRY(tm.stop(c, &ms));
    const int converged = hstate[0] && !hstate[2];
    const int iters = hstate[0] ? hstate[1] : p.max_iter;
    std::vector<double> hh((size_t)std::max(iters, 1), 0.0);
    if (iters > 0)
        SENK_HIP(hipMemcpy(hh.data(), dhist, sizeof(double) * std::min(iters, hcap), hipMemcpyDeviceToHost));
    if (hist)
        for (int k = 0; k < std::min(iters, hist_cap); k++) hist[k] = hh[k];
    res->iters = iters;
    res->converged = converged;
    res->restarts = 0;
    res->implicit_relres = iters > 0 ? hh[iters - 1] : 0;
    res->t_solve_ms = ms;
    res->status = hstate[2] ? SENK_ERR_BREAKDOWN : SENK_OK;
    if (p.verbose) {
        if (converged) {
            std::printf("I iter %d\n", iters);
            std::printf("# res %e\n", res->implicit_relres);
        } else {
            std::printf("# iter %d (max)\n", iters);
            std::printf("# res %e\n", res->implicit_relres);
        }
        std::fflush(stdout);
    }
    return SENK_OK;
}

template <typename T>
int krylov_solve(Ctx *c, const MatView &A, const PrecView &M, const T *b, T *x, const SolveParams &p,
                 senk_result *res, double *hist, int hist_cap) {
    std::memset(res, 0, sizeof(*res));
    if (A.N <= 0 || p.m <= 0 || p.m > 512) return SENK_ERR_ARG;
    switch (p.solver) {
        case SENK_GMRES:
        case SENK_FGMRES: return gmres<T>(c, A, M, b, x, p, res, hist, hist_cap);
        case SENK_GCR:
        case SENK_FGCR: return gcr<T>(c, A, M, b, x, p, res, hist, hist_cap);
        case SENK_BICGSTAB: return bicgstab<T>(c, A, M, b, x, p, res, hist, hist_cap);
        case SENK_CG: return cg<T>(c, A, M, b, x, p, res, hist, hist_cap);
        default: return SENK_ERR_ARG;
    }
}
template int krylov_solve<double>(Ctx *, const MatView &, const PrecView &, const double *, double *,
                                  const SolveParams &, senk_result *, double *, int);
template int krylov_solve<float>(Ctx *, const MatView &, const PrecView &, const float *, float *,
                                 const SolveParams &, senk_result *, double *, int);

// ================================================================= mixed precision
namespace {
// x += ||r|| * d  (FP64 x, FP32 d; ||r||^2 on the device)
__global__ __launch_bounds__(kBlock) void k_ir_update(int N, const float *__restrict__ d, const double *nrm2,
                                                      double *x) {
    const double s = sqrt(*nrm2);
    SENK_FOR4(i0, N) {
        float dv[4];
        double xv[4];
        get4(d, i0, N, dv);
        get4(x, i0, N, xv);
#pragma unroll
        for (int u = 0; u < 4; u++) xv[u] += s * (double)dv[u];
        put4(x, i0, N, xv);
    }
}
}  // namespace

int true_relres(Ctx *c, const MatView &A, const double *b, const double *x, double nrm_b, double *out) {
    void *wsp;
    SENK_TRY(ctx_work(c, bytes_of<double>(A.N) + bytes_of<double>(4), &wsp));
    Carver cv{(char *)wsp, 0, 0};
    double *r = cv.take<double>(A.N);
    double *d = cv.take<double>(4);
    int rg = 0;
    SENK_TRY(residual<double>(c, A, x, b, r, &rg));
    k_norm_fin<<<1, kBlock, 0, c->stream>>>(red_slot(c, kSlotRes), rg, d, nullptr);
    SENK_LAUNCH_CHECK();
    double h = 0;
    SENK_HIP(hipMemcpyAsync(&h, d, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SENK_HIP(hipStreamSynchronize(c->stream));
    *out = std::sqrt(h) / nrm_b;
    return SENK_OK;
}

// FP64 iterative refinement around an FP32 Krylov solve (GMRES-IR): the FP64 residual
// r = b - A x drives the stop test; the inner solve gets r/||r|| in FP32 and its
// correction is added back in FP64.
namespace {
// q -= sum_k coef[k] Q_k and z -= sum_k coef[k] Z_k (vectors k at stride ld)
__global__ __launch_bounds__(kBlock) void k_sub_lincomb2(int N, int K, const double *__restrict__ coef,
                                                         const double *__restrict__ Q, const double *__restrict__ Z,
                                                         size_t ld, double *q, double *z) {
    SENK_FOR4(i0, N) {
        double sq[4], sz[4], t[4];
        get4(q, i0, N, sq);
        get4(z, i0, N, sz);
        for (int k = 0; k < K; k++) {
            const double a = coef[k];
            get4(Q + (size_t)k * ld, i0, N, t);
#pragma unroll
            for (int u = 0; u < 4; u++) sq[u] -= a * t[u];
            get4(Z + (size_t)k * ld, i0, N, t);
#pragma unroll
            for (int u = 0; u < 4; u++) sz[u] -= a * t[u];
        }
        put4(q, i0, N, sq);
        put4(z, i0, N, sz);
    }
}
// s = {(q,q), (r,q)}: alpha = s1/s0, x += alpha z, r -= alpha q, then q and z are scaled
// by 1/||q|| so that the stored directions satisfy A z = q with ||q|| = 1.
__global__ __launch_bounds__(kBlock) void k_gcr_outer_step(int N, const double *__restrict__ s, double *q,
                                                           double *z, double *x, double *r) {
    const double alpha = s[1] / s[0], inv = 1.0 / sqrt(s[0]);
    SENK_FOR4(i0, N) {
        double qv[4], zv[4], t[4];
        get4(q, i0, N, qv);
        get4(z, i0, N, zv);
        get4(x, i0, N, t);
#pragma unroll
        for (int u = 0; u < 4; u++) t[u] += alpha * zv[u];
        put4(x, i0, N, t);
        get4(r, i0, N, t);
#pragma unroll
        for (int u = 0; u < 4; u++) t[u] -= alpha * qv[u];
        put4(r, i0, N, t);
#pragma unroll
        for (int u = 0; u < 4; u++) {
            qv[u] *= inv;
            zv[u] *= inv;
        }
        put4(q, i0, N, qv);
        put4(z, i0, N, zv);
    }
}
__global__ void k_inv_sqrt(const double *s, double *inv) {
    if (threadIdx.x == 0) *inv = 1.0 / sqrt(*s);
}

// The learned correction is switched off for the rest of a solve when an outer step
// (or a restart cycle) fails to reduce the FP64 residual by this factor.
constexpr double kStagnation = 0.5;
void monitor_learned(const PrecView &M, double rho, double rho_prev) {
    if (M.learned && !M.learned->off && rho_prev > 0 && !(rho <= kStagnation * rho_prev)) M.learned->off = true;
}

SolveParams inner_params(const senk_solve_opts &o) {
    SolveParams ip;
    ip.solver = o.solver;
    ip.m = o.m;
    ip.eps = o.inner_eps;
    ip.nrm_b = 1.0;  // the inner right-hand side is normalised
    ip.ortho = o.ortho;
    ip.check_every = 1;
    ip.max_iter = o.inner_max;
    ip.outer = (o.inner_max + o.m - 1) / o.m;
    ip.split = 1;  // both callers start the inner solve from d = 0
    return ip;
}

// The inner solve of an outer step that starts at residual norm rho, against the outer
// tolerance tol (same units).  An inner solve that leaves the fraction e of its
// right-hand side's residual takes the outer residual to at most e rho (the refinement's
// x += d, or the flexible GCR step, which minimises over the new direction), so the last
// outer step needs no smaller e than tol / rho: the inner tolerance is max(inner_eps,
// kInnerRelax * tol / rho), which only moves in the last outer steps.  An inner solve
// that falls short costs one more outer step at a looser tolerance still.  (128^3
// variable-coefficient system, nested FP32 split-BiCGStab(60): factor 0 / 0.1 / 0.3 / 0.5,
// 233 / 219 / 203 / 198 inner iterations, 32.7 / 31.7 / 28.8 / 28.2 ms; nested split
// CG(100): 304 inner iterations at every factor.)
constexpr double kInnerRelax = 0.5;
SolveParams relaxed_inner(const SolveParams &ip, double tol, double rho) {
    SolveParams q = ip;
    if (rho > 0 && std::isfinite(rho)) q.eps = std::max(ip.eps, kInnerRelax * tol / rho);
    return q;
}

// The learned correction of a mixed form: with a flexible inner solver (FGCR, FGMRES) it
// is added to every ILU application of the inner solve; with any other inner solver it is
// added once to the inner solve's result d (the refinement's correction, or the outer
// flexible GCR's preconditioner application), from the inner right-hand side r and d.
// Returns the preconditioner the inner solve uses.
PrecView inner_prec(const PrecView &M, const senk_solve_opts &o, bool *learned_after) {
    PrecView Mi = M;
    *learned_after = M.learned && o.solver != SENK_FGCR && o.solver != SENK_FGMRES;
    if (*learned_after) Mi.learned = nullptr;
    return Mi;
}

int read_scalar(Ctx *c, const double *d, double *h) {
    SENK_HIP(hipMemcpyAsync(h, d, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SENK_HIP(hipStreamSynchronize(c->stream));
    return SENK_OK;
}
}  // namespace

// Iterative refinement: r = b - A x in FP64, d = inner FP32 solve of A d = r/||r||,
// x += ||r|| d, until ||r|| <= eps ||b||.
int ir_solve(Ctx *c, const MatView &A, const PrecView &M, const double *b, double *x, const senk_solve_opts &o,
             double nrm_b, senk_result *res, double *hist, int hist_cap) {
    std::memset(res, 0, sizeof(*res));
    const int N = A.N;
    if (!A.v32) return SENK_ERR_ARG;
    // the inner solver carves its workspace from the context arena; the refinement
    // vectors live in the outer arena
    const size_t need = bytes_of<double>(N) + 2 * bytes_of<float>(N) + bytes_of<double>(4);
    void *wsp;
    SENK_TRY(ctx_outer(c, need, &wsp));
    Carver cv{(char *)wsp, 0, need};
    double *r64 = cv.take<double>(N);
    float *r32 = cv.take<float>(N), *d32 = cv.take<float>(N);
    double *dsc = cv.take<double>(4);
    SolveTimer tm;
    SENK_TRY(tm.start(c, true));
    const SolveParams ip = inner_params(o);
    bool learned_after = false;
    const PrecView Mi = inner_prec(M, o, &learned_after);
    const int max_steps = std::max(1, (o.max_iter + o.inner_max - 1) / std::max(1, o.inner_max)) + 8;
    int steps = 0, inner_total = 0, converged = 0, status = SENK_OK;
    double relres = 0;
    for (;;) {
        int rg = 0;
        SENK_TRY(residual<double>(c, A, x, b, r64, &rg));
        k_norm_fin<<<1, kBlock, 0, c->stream>>>(red_slot(c, kSlotRes), rg, dsc, dsc + 1);
        SENK_LAUNCH_CHECK();
        double h2 = 0;
        SENK_TRY(read_scalar(c, dsc, &h2));
        const double prev = relres;
        relres = std::sqrt(h2) / nrm_b;
        if (steps > 0) monitor_learned(M, relres, prev);
        if (hist && steps < hist_cap) hist[steps] = relres;
        if (o.verbose) {
            std::printf("IR %d %e\n", steps, relres);
            std::fflush(stdout);
        }
        if (!std::isfinite(relres)) {
            status = SENK_ERR_BREAKDOWN;
            break;
        }
        if (relres <= o.eps) {
            converged = 1;
            break;
        }
        if (steps >= max_steps || inner_total >= o.max_iter) break;
        SENK_TRY(launch_convert<double, float>(c, N, r64, r32, dsc + 1));
        SENK_HIP(hipMemsetAsync(d32, 0, sizeof(float) * N, c->stream));
        senk_result ir;
        SENK_TRY(krylov_solve<float>(c, A, Mi, r32, d32, relaxed_inner(ip, o.eps, relres), &ir, nullptr, 0));
        inner_total += ir.iters;
        if (learned_after && !M.learned->off) SENK_TRY(learned_apply_f32(c, M, r32, nullptr, d32, nullptr));
        k_ir_update<<<ew_grid(N), kBlock, 0, c->stream>>>(N, d32, dsc, x);
        SENK_LAUNCH_CHECK();
        steps++;
    }
    float ms = 0;
    SENK_TRY(tm.stop(c, &ms));
    res->iters = inner_total;
    res->inner_iters = inner_total;
    res->restarts = steps;
    res->converged = converged;
    res->implicit_relres = relres;
    res->true_relres = relres;
    res->t_solve_ms = ms;
    res->status = status;
    return SENK_OK;
}

// Nested Krylov: FP64 flexible GCR(outer_m) on A x = b whose preconditioner is an
// inner FP32 Krylov solve (o.solver, restart o.m, o.inner_max iterations or relative
// tolerance o.inner_eps) with the FP32 matrix and factors.  Outer step j:
//   z_j = inner(r / ||r||),  q_j = A z_j (FP64),
//   q_j, z_j orthogonalised against q_0..q_{j-1} (classical Gram-Schmidt, twice),
//   alpha = (r, q_j)/(q_j, q_j),  x += alpha z_j,  r -= alpha q_j.
// The outer basis keeps what every inner solve found, so unlike refinement nothing is
// discarded between inner solves.  The FP64 residual is recomputed at every restart
// and before convergence is reported.
int nested_solve(Ctx *c, const MatView &A, const PrecView &M, const double *b, double *x,
                 const senk_solve_opts &o, double nrm_b, senk_result *res, double *hist, int hist_cap,
                 NestedRecord *rec) {
    std::memset(res, 0, sizeof(*res));
    const int N = A.N;
    const int mo = o.outer_m > 0 ? o.outer_m : 20;
    const size_t ld = pad4(N);
    if (!A.v32 || mo > kRedMaxK) return SENK_ERR_ARG;
    const size_t need = 2 * bytes_of<double>((size_t)mo * ld) + bytes_of<double>(N) + 2 * bytes_of<float>(N) +
                        bytes_of<double>(kRedMaxK) + bytes_of<double>(8);
    void *wsp;
    SENK_TRY(ctx_outer(c, need, &wsp));
    Carver cv{(char *)wsp, 0, need};
    double *Q = cv.take<double>((size_t)mo * ld), *Z = cv.take<double>((size_t)mo * ld);
    double *r = cv.take<double>(N);
    float *r32 = cv.take<float>(N), *d32 = cv.take<float>(N);
    double *beta = cv.take<double>(kRedMaxK), *sc = cv.take<double>(8);
    SolveTimer tm;
    SENK_TRY(tm.start(c, true));
    const SolveParams ip = inner_params(o);
    bool learned_after = false;
    const PrecView Mi = inner_prec(M, o, &learned_after);
    const double tol = o.eps * nrm_b;
    int inner_total = 0, outer_its = 0, cycles = 0, converged = 0, status = SENK_OK, hn = 0;
    double rho = 0;
    for (;;) {
        int rg = 0;
        SENK_TRY(residual<double>(c, A, x, b, r, &rg));
        k_norm_fin<<<1, kBlock, 0, c->stream>>>(red_slot(c, kSlotRes), rg, sc, sc + 1);
        SENK_LAUNCH_CHECK();
        double h2 = 0;
        SENK_TRY(read_scalar(c, sc, &h2));
        rho = std::sqrt(h2);
        if (hist && hn < hist_cap) hist[hn++] = rho / nrm_b;
        if (!std::isfinite(rho)) {
            status = SENK_ERR_BREAKDOWN;
            break;
        }
        if (rho <= tol) {
            converged = 1;
            break;
        }
        if (inner_total >= o.max_iter) break;
        cycles++;
        for (int j = 0; j < mo; j++) {
            double *q = Q + (size_t)j * ld, *z = Z + (size_t)j * ld;
            SENK_TRY(launch_convert<double, float>(c, N, r, r32, sc + 1));
            SENK_HIP(hipMemsetAsync(d32, 0, sizeof(float) * N, c->stream));
            senk_result ir;
            SENK_TRY(krylov_solve<float>(c, A, Mi, r32, d32, relaxed_inner(ip, tol, rho), &ir, nullptr, 0));
            inner_total += ir.iters;
            if (rec && rec->count < rec->cap) {
                SENK_TRY(launch_copy<float>(c, N, r32, rec->R + (size_t)rec->count * N));
                SENK_TRY(launch_copy<float>(c, N, d32, rec->Z + (size_t)rec->count * N));
                rec->count++;
            }
            if (learned_after && !M.learned->off) SENK_TRY(learned_apply_f32(c, M, r32, nullptr, d32, nullptr));
            SENK_TRY(launch_convert<float, double>(c, N, d32, z, nullptr));
            SENK_TRY(spmv<double>(c, A, z, q));
            for (int pass = 0; pass < 2 && j > 0; pass++) {
                SENK_TRY(launch_multidot<double>(c, N, j, Q, ld, q, beta));
                k_sub_lincomb2<<<ew_grid(N), kBlock, 0, c->stream>>>(N, j, beta, Q, Z, ld, q, z);
                SENK_LAUNCH_CHECK();
            }
            SENK_TRY(launch_multidot<double>(c, N, 1, q, (size_t)N, q, sc + 2));
            SENK_TRY(launch_multidot<double>(c, N, 1, q, (size_t)N, r, sc + 3));
            k_gcr_outer_step<<<ew_grid(N), kBlock, 0, c->stream>>>(N, sc + 2, q, z, x, r);
            SENK_LAUNCH_CHECK();
            SENK_TRY(launch_multidot<double>(c, N, 1, r, (size_t)N, r, sc));
            k_inv_sqrt<<<1, 64, 0, c->stream>>>(sc, sc + 1);
            SENK_LAUNCH_CHECK();
            SENK_TRY(read_scalar(c, sc, &h2));
            const double rho_prev = rho;
            rho = std::sqrt(h2);
            monitor_learned(M, rho, rho_prev);
            outer_its++;
            if (hist && hn < hist_cap) hist[hn++] = rho / nrm_b;
            if (o.verbose) {
                std::printf("outer %d inner %d %e\n", outer_its, inner_total, rho / nrm_b);
                std::fflush(stdout);
            }
            if (!std::isfinite(rho) || rho <= tol || inner_total >= o.max_iter) break;
        }
        if (!std::isfinite(rho)) {
            status = SENK_ERR_BREAKDOWN;
            break;
        }
    }
    float ms = 0;
    SENK_TRY(tm.stop(c, &ms));
    res->iters = inner_total;
    res->inner_iters = inner_total;
    res->restarts = outer_its;
    res->converged = converged;
    res->implicit_relres = rho / nrm_b;
    res->true_relres = rho / nrm_b;
    res->t_solve_ms = ms;
    res->status = status;
    (void)cycles;
    return SENK_OK;
}

// Residual/error pairs of the plain FP64 GMRES(m): after each restart cycle from
// x0 = 0, r = b - A x (and x* - x) become row k of R (E).
int collect_pairs(Ctx *c, const MatView &A, const PrecView &M, const double *b, const double *xstar, int m,
                  int ncycles, double *R, double *E, int *count) {
    const int N = A.N;
    void *wsp;
    SENK_TRY(ctx_outer(c, bytes_of<double>(N) + bytes_of<double>(4), &wsp));
    Carver cv{(char *)wsp, 0, 0};
    double *x = cv.take<double>(N), *d = cv.take<double>(4);
    SENK_HIP(hipMemsetAsync(x, 0, sizeof(double) * N, c->stream));
    int rg = 0;
    SENK_TRY(residual<double>(c, A, x, b, R, &rg));
    k_norm_fin<<<1, kBlock, 0, c->stream>>>(red_slot(c, kSlotRes), rg, d, nullptr);
    SENK_LAUNCH_CHECK();
    double h = 0;
    SENK_TRY(read_scalar(c, d, &h));
    const double nrm_b = std::sqrt(h);
    SolveParams p;
    p.solver = SENK_GMRES;
    p.m = m;
    p.outer = 1;
    p.eps = 0;  // run whole cycles
    p.nrm_b = nrm_b;
    int k = 0;
    while (k < ncycles) {
        senk_result res;
        SENK_TRY(krylov_solve<double>(c, A, M, b, x, p, &res, nullptr, 0));
        double *rk = R + (size_t)k * N;
        SENK_TRY(residual<double>(c, A, x, b, rk, &rg));
        k_norm_fin<<<1, kBlock, 0, c->stream>>>(red_slot(c, kSlotRes), rg, d, nullptr);
        SENK_LAUNCH_CHECK();
        if (E && xstar) SENK_TRY(launch_axpyz<double>(c, N, -1.0, x, xstar, E + (size_t)k * N));
        SENK_TRY(read_scalar(c, d, &h));
        k++;
        if (!(std::sqrt(h) > 1e-13 * nrm_b)) break;
    }
    *count = k;
    return SENK_OK;
}

}  // namespace senk_amd
