// Internal declarations shared by the HIP kernels, the host setup code and the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "senk_amd.h"

namespace senk_amd {

// ----------------------------------------------------------------- errors
void record_hip_error(hipError_t e, const char *what, const char *file, int line);
void record_error(const std::string &msg);

#define SENK_HIP(...)                                                                   \
    do {                                                                                \
        hipError_t e_ = (__VA_ARGS__);                                                  \
        if (e_ != hipSuccess) {                                                         \
            ::senk_amd::record_hip_error(e_, #__VA_ARGS__, __FILE__, __LINE__);         \
            return SENK_ERR_HIP;                                                        \
        }                                                                               \
    } while (0)
#define SENK_TRY(...)                                                                   \
    do {                                                                                \
        int s_ = (__VA_ARGS__);                                                         \
        if (s_ != SENK_OK) return s_;                                                   \
    } while (0)
#define SENK_LAUNCH_CHECK() SENK_HIP(hipGetLastError())

// ----------------------------------------------------------------- device buffers
// A grow-only device arena owned by a context: solver work vectors are carved out of
// it, so repeated solves allocate nothing after the first.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

// Reduction scratch: one partial sum per workgroup and dot product; the consuming
// kernel adds them in a fixed order.
constexpr int kRedMaxBlocks = 1024;  // workgroups of a reducing pass
constexpr int kRedMaxK = 40;          // simultaneous dot products per pass
constexpr int kRedSlots = 8;          // independent partial buffers

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    double *red_partials = nullptr;   // kRedSlots * kRedMaxK * kRedMaxBlocks
    double *dscalars = nullptr;       // 64 general device scalars
    double *hpinned = nullptr;        // 64 pinned host doubles
    int *hpinned_i = nullptr;         // 64 pinned host ints
    DevBuf work;                      // solver arena
    DevBuf outer;                     // arena of the FP64 outer loop around an inner solve
    int num_cus = 256;
    // events that solves reuse (created on first use): [8, 10) inner solve timer, [10, 12)
    // outer solve timer
    hipEvent_t ev[12] = {};
    // coherent pinned ring the solvers' last kernel of an iteration writes itself (host
    // and device addresses), 8 slots of 4 ints: {sequence number, stop flag} in the first 8
    // bytes of each
    int *dring_h = nullptr, *dring_d = nullptr;
    unsigned dseq = 0;  // last sequence number handed out for the ring (runs on across solves)
};
// Creates the context's reusable events and pinned ring on first use.
int ctx_events(Ctx *c);

int ctx_work(Ctx *c, size_t bytes, void **out);
int ctx_outer(Ctx *c, size_t bytes, void **out);
inline double *red_slot(Ctx *c, int slot) { return c->red_partials + (size_t)slot * kRedMaxK * kRedMaxBlocks; }

// ----------------------------------------------------------------- schedules
// Blocks of `bsize` consecutive rows run in `nlev` launches; launch l covers
// blocks blk[lev_ptr[l] .. lev_ptr[l+1]) (blk == nullptr: identity).  For an upper
// solve the launches run from the last to the first.
struct Schedule {
    int bsize = 1;
    int nlev = 0;
    std::vector<int> lev_ptr;  // host
    int *dblk = nullptr;       // device, optional
};

// A triangular factor rearranged for the scheduled solve (trsv_sell.hip): per level,
// slices of 64 rows; per row the external entries (value + column) and the in-block
// entries (value; column = bit of `mask`), each group stored packed and in the order
// the solve consumes it.  Built once per pattern, filled per precision.
// Launch modes of the layout solve: one launch per level, or one launch per triangle
// whose gathers wait for the rows they read (sentinel-initialised output), with every
// wave resident (cooperative) or with workgroups drawing tickets in arrival order.
constexpr int kTrsvLevels = 0, kTrsvSweep = 1, kTrsvTicket = 2, kTrsvXcd = 3;
struct TriLayout {
    int bsize = 0;
    bool upper = false;
    int n = 0;                   // rows
    int nslices = 0;
    int nsteps = 0;              // non-empty levels, in solve order
    std::vector<int> step_slice; // host, nsteps + 1: first slice of each step
    long long n_ext = 0, n_int = 0;
    int max_ext = 0;             // largest external entry count of a row
    int4 *sinfo = nullptr;       // per slice: external offset, in-block offset, first block entry, blocks
    void *rw = nullptr;          // per slot (slice*64 + lane): external count and in-block mask
    int *next = nullptr;         // bsize 64 only: per slot external count
    int *scol = nullptr;         // n_ext external columns
    bool cols_filled = false;
    double *ve64 = nullptr, *vi64 = nullptr, *d64 = nullptr;  // values; d = inverted diagonal (U)
    float *ve32 = nullptr, *vi32 = nullptr, *d32 = nullptr;
    unsigned *sync = nullptr;    // sweep time-out flag, spares
    void *mid64 = nullptr, *mid32 = nullptr;  // coherent output vector of the sweep
    // FP16 copy: values scaled by 2^-e per row (e in sx16), FP32 inverted diagonal (U)
    void *ve16 = nullptr, *vi16 = nullptr;
    float *d16 = nullptr;
    signed char *sx16 = nullptr;
    int mode = kTrsvXcd;
    unsigned ticket_base = 0;  // ticketed sweep: workgroups launched so far (mod 2^32)
    // per-XCD ticket streams (mode 3): stream x holds the x-th eighth of every level's
    // slices, each level padded to the same number of 8-slice tickets in every stream;
    // xwg tickets (workgroups) per stream.  tord: slices numbered in ticket order (wave w of ticket j
    // of stream x solves slice 64 j + 8 x + w; bands padded with empty slices); else
    // slice order and the table xslice[x * 8 * xwg + 8 j + w] (-1: none)
    bool tord = false;
    int *xslice = nullptr;
    int xwg = 0;
};
int tri_layout_check(Ctx *c, TriLayout *L);
int tri_layout_build(Ctx *c, const Schedule &s, bool upper, int N, const int *rptr, const int *cind,
                     TriLayout *L);
template <typename T>
int tri_layout_fill(Ctx *c, const Schedule &s, TriLayout *L, const int *rptr, const int *cind, const T *val);
void tri_layout_free(TriLayout *L);
// ILU application out = U^-1 L^-1 (in_scale * in) with the two layouts; in may equal out.
// f16 (T = float only): the FP16 copies of the factor values.
template <typename T>
int launch_ilu_layout(Ctx *c, const Schedule &ls, TriLayout &lt, const Schedule &us, TriLayout &ut,
                      const T *in, T *out, const double *in_scale, const int *stop, bool f16 = false);
int tri_layout_fill_f16(Ctx *c, const Schedule &s, TriLayout *L, const int *rptr, const int *cind, const double *val);
// out = self^-1 in for one triangle alone (self L or U, other the opposite triangle of
// the same factorization); in may equal out.  Calls on one factorization must alternate
// L, U, L, ... (an ILU application counts as L then U).
// With both epilogue vectors (single-launch modes only; *fused says whether they were
// used) a U sweep also writes ep_b = in - ep_a * out, and an L sweep writes
// out = ep_a * ep_b + L^-1 in (ep_b read; not the same vector as out; it may be the U
// sweep's own vector, each row read before this sweep returns it to the sentinel).
// A U solve may first update its right-hand side (pre.u == in, non-null): with
// a = pre.ab[0], b = pre.ab[1], per row yd += a u, r = r - a q, u = r + b u, and the
// solve uses the new u (the split CG's vector step folded into the next product).  The
// single-launch modes do it in the sweep as each row's right-hand side is read; the
// level-by-level mode in a pass before the solve.  Skipped, like the solve, when *stop is
// set.
template <typename T>
struct SweepPre {
    T *u = nullptr;
    T *r = nullptr;
    const T *q = nullptr;
    T *yd = nullptr;
    const T *ab = nullptr;
};
// out may be null where the triangle is solved by a single-launch sweep: the result is
// then only in the sweep's own vector, tri_single_sweep_vector(self, other) (null: the
// per-level mode, which has none), valid until the other triangle's next solve returns it
// to the sentinel.
template <typename T>
int launch_tri_single(Ctx *c, const Schedule &s, TriLayout &self, TriLayout &other, const T *in, T *out,
                      const int *stop, const T *ep_a = nullptr, T *ep_b = nullptr, bool *fused = nullptr,
                      const SweepPre<T> &pre = SweepPre<T>{});
template <typename T>
T *tri_single_sweep_vector(const TriLayout &self, const TriLayout &other);

// ----------------------------------------------------------------- launchers (kernels.hip)
template <typename T>
int launch_spmv_csr(Ctx *c, int N, const int *rptr, const int *cind, const T *val, const T *x, T *y,
                    double avg_row, const int *stop = nullptr);
int launch_spmv_csr_diag_f64(Ctx *c, int N, const int *rptr, const int *cind, const double *val,
                             const double *diag, const double *x, double *y);
template <typename T>
int launch_spmv_bcsr(Ctx *c, int N, int bnl, int bnw, const int *brptr, const int *bcind, const T *bval,
                     const T *x, T *y);
template <typename T>
int launch_spmv_sell(Ctx *c, int N, int C, int nslices, const int *wid, const int *cind, const T *val,
                     const int *perm, const T *x, T *y, bool last_slice_short, const int *stop = nullptr,
                     double avg_width = 0, const int *order = nullptr, bool nontemporal = false,
                     long long slots = -1);  // stored slots (wid[nslices] * C) when known

template <typename T>
int launch_sptrsv(Ctx *c, const Schedule &s, bool upper, const int *rptr, const int *cind, const T *val,
                  const T *x, T *y, const double *in_scale, const int *stop_flag);
template <typename T>
int launch_sptrsv_bcsr(Ctx *c, const Schedule &s, bool upper, int bnl, int bnw, const int *brptr,
                       const int *bcind, const T *bval, const T *x, T *y);
template <typename T>
int launch_ilu0(Ctx *c, const Schedule &s, int N, const int *rptr, const int *cind, const int *dpos, T *val,
                int *derr);
int launch_find_diag(Ctx *c, int N, const int *rptr, const int *cind, int *dpos);
// d = diag(U) (from U's inverted diagonal, first in each row), e = 2 d - diag(A), FP32
int launch_ic_upper(Ctx *c, int lnnz, const int *umap, const int *lcol, const double *f, const int *dpos,
                    const double *l, double *u);
int launch_dilu_diag(Ctx *c, int N, const int *urptr, const double *uval, const int *arptr, const int *acind,
                     const double *aval, float *d, float *e);
template <typename T>
int launch_split_ldu(Ctx *c, int N, const int *rptr, const int *cind, const T *val, const int *lrptr,
                     T *lval, int *lcind, const int *urptr, T *uval, int *ucind);

// blas1
template <typename T>
int launch_axpby(Ctx *c, int N, double a, const T *x, double b, T *y);       // y = a x + b y
template <typename T>
int launch_axpyz(Ctx *c, int N, double a, const T *x, const T *y, T *z);     // z = a x + y
template <typename T>
int launch_scal(Ctx *c, int N, double a, T *x);
template <typename T>
int launch_copy(Ctx *c, int N, const T *x, T *y);
template <typename T>
int launch_hadprod(Ctx *c, int N, const T *x, T *y);                        // y *= x
template <typename T>
int launch_haddiv(Ctx *c, int N, const T *x, T *y);                         // y /= x
template <typename TI, typename TO>
int launch_convert(Ctx *c, int N, const TI *in, TO *out, const double *dscale);
// result[k] = sum_i x_k[i]*y[i] for k < K (x_k = xs + k*ldx), FP64 accumulation,
// deterministic.  result is a device pointer.
template <typename T>
int launch_multidot(Ctx *c, int N, int K, const T *xs, size_t ldx, const T *y, double *result);
int launch_scaling_f64(Ctx *c, int N, const int *rptr, const int *cind, double *val, const double *ld,
                       const double *rd);

// ----------------------------------------------------------------- host setup (host_setup.cpp)
namespace host {
int abmc(const int *cind, const int *rptr, int N, int bsize, bool isSym, bool connect, std::vector<int> &size_color,
         std::vector<int> &LP, std::vector<int> &RP);
int amc(const int *cind, const int *rptr, int N, bool isSym, std::vector<int> &size_color, std::vector<int> &LP,
        std::vector<int> &RP);
// Dependency levels of the blocks of bsize rows of a triangular pattern (lower: columns
// < row; upper: columns > row).  Returns levels in execution order.
int levels(const int *rptr, const int *cind, int N, int bsize, bool upper, std::vector<int> &lev_ptr,
           std::vector<int> &blk);
// Checks that no entry couples two different blocks of the same launch.
bool schedule_ok(const int *rptr, const int *cind, int N, const Schedule &s, const std::vector<int> &blk,
                 bool lower_part, bool upper_part);
}  // namespace host

}  // namespace senk_amd
