// C ABI glue: contexts, device memory, kernel entry points with the reference's
// argument order, device ILU factorization, the solver twins and the handle API.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <numeric>
#include <string>
#include <vector>

#include "internal.hpp"
#include "solvers.hpp"

struct senk_ctx_s {
    senk_amd::Ctx c;
    std::map<std::pair<const void *, int>, int> nnz_cache;  // rptr -> nnz (SpMV lane choice)
};
struct senk_trsv_plan_s {
    senk_amd::Ctx *c = nullptr;
    senk_amd::Schedule s;
};
struct senk_mat_s {
    senk_ctx ctx = nullptr;
    int N = 0, nnz = 0;
    int *rptr = nullptr, *cind = nullptr;
    double *v64 = nullptr;
    float *v32 = nullptr;
    bool has_sell = false;
    senk_sellcs sell{};
    int *sell_order[2] = {nullptr, nullptr};  // device: slice orders by 8 and by 64 regions of every color
    senk_amd::MatView view;
};
struct senk_prec_s {
    senk_ctx ctx = nullptr;
    senk_amd::PrecView view;
    senk_amd::TriLayout lt, ut;
    float *dilu_d = nullptr, *dilu_e = nullptr;  // D-ILU vectors (owned; null: not D-ILU)
    senk_gnn gnn = nullptr;  // attached learned correction (used with use_learned)
    int lnnz = 0, unnz = 0;
    std::vector<void *> owned;
    std::vector<int> hlrptr, hurptr;
};

namespace senk_amd {
namespace {
thread_local std::string g_last_error;
}
void record_hip_error(hipError_t e, const char *what, const char *file, int line) {
    char buf[512];
    std::snprintf(buf, sizeof(buf), "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    g_last_error = buf;
}
void record_error(const std::string &msg) { g_last_error = msg; }

static int arena_get(DevBuf &a, size_t bytes, void **out) {
    if (a.bytes < bytes) {
        if (a.p) SENK_HIP(hipFree(a.p));
        a.p = nullptr;
        a.bytes = 0;
        SENK_HIP(hipMalloc(&a.p, bytes));
        a.bytes = bytes;
    }
    *out = a.p;
    return SENK_OK;
}
int ctx_work(Ctx *c, size_t bytes, void **out) { return arena_get(c->work, bytes, out); }
int ctx_events(Ctx *c) {
    if (c->dring_h) return SENK_OK;
    for (int i = 8; i < 12; i++) SENK_HIP(hipEventCreateWithFlags(&c->ev[i], hipEventDefault));
    SENK_HIP(hipHostMalloc((void **)&c->dring_h, sizeof(int) * 4 * 8, hipHostMallocCoherent));
    std::memset(c->dring_h, 0, sizeof(int) * 4 * 8);
    SENK_HIP(hipHostGetDevicePointer((void **)&c->dring_d, c->dring_h, 0));
    return SENK_OK;
}
int ctx_outer(Ctx *c, size_t bytes, void **out) { return arena_get(c->outer, bytes, out); }

namespace {
int nnz_of(senk_ctx ctx, const int *drptr, int N, int *nnz) {
    auto key = std::make_pair((const void *)drptr, N);
    auto it = ctx->nnz_cache.find(key);
    if (it != ctx->nnz_cache.end()) {
        *nnz = it->second;
        return SENK_OK;
    }
    int h[2] = {0, 0};
    SENK_HIP(hipMemcpyAsync(&h[0], drptr, sizeof(int), hipMemcpyDeviceToHost, ctx->c.stream));
    SENK_HIP(hipMemcpyAsync(&h[1], drptr + N, sizeof(int), hipMemcpyDeviceToHost, ctx->c.stream));
    SENK_HIP(hipStreamSynchronize(ctx->c.stream));
    *nnz = h[1] - h[0];
    if (ctx->nnz_cache.size() > 256) ctx->nnz_cache.clear();
    ctx->nnz_cache[key] = *nnz;
    return SENK_OK;
}
double avg_row(int nnz, int N) { return N > 0 ? (double)nnz / N : 0.0; }

// True when ILU(0) on this pattern changes no off-diagonal entry: no row i has an entry
// k < i whose row holds an entry j > k, j != i, that row i holds as well (the IKJ
// elimination of senk_matrix.hpp updates exactly those), so the factors' off-diagonal
// entries are A's own and only the pivots change (a triangle-free graph, e.g. a 5- or
// 7-point stencil in any ordering).
// A symmetric in pattern, and in values to a relative tolerance rtol (0: bit for bit):
// |a_ij - a_ji| <= rtol * max(|a_ij|, |a_ji|).  A symmetric matrix that went through
// Jacobi scaling (val = ld_i * a_ij * rd_j, senk_matrix.hpp:406-416) keeps its symmetry only
// to the last bit or two of each product.
constexpr double kSymTol = 16 * 2.220446049250313e-16;  // a few units in the last place
bool symmetric_csr(int N, const int *rp, const int *ci, const double *v, double rtol) {
    std::vector<int> tp((size_t)N + 1, 0);
    const int nnz = rp[N];
    for (int j = 0; j < nnz; j++) {
        if (ci[j] < 0 || ci[j] >= N) return false;
        tp[ci[j] + 1]++;
    }
    for (int i = 0; i < N; i++) {
        tp[i + 1] += tp[i];
        if (tp[i + 1] - tp[i] != rp[i + 1] - rp[i]) return false;
    }
    std::vector<std::pair<int, double>> t((size_t)nnz);
    std::vector<int> fill(tp.begin(), tp.end() - 1);
    for (int i = 0; i < N; i++)
        for (int j = rp[i]; j < rp[i + 1]; j++) t[fill[ci[j]]++] = {i, v[j]};
    std::vector<std::pair<int, double>> row;
    for (int i = 0; i < N; i++) {
        row.clear();
        for (int j = rp[i]; j < rp[i + 1]; j++) row.push_back({ci[j], v[j]});
        std::sort(row.begin(), row.end());
        // column i of A, rows ascending by construction
        for (int k = 0; k < (int)row.size(); k++) {
            const double a = row[k].second, b = t[tp[i] + k].second;
            if (row[k].first != t[tp[i] + k].first) return false;
            if (!(a == b) && !(std::fabs(a - b) <= rtol * std::max(std::fabs(a), std::fabs(b)))) return false;
        }
    }
    return true;
}

// Overwrite every strictly upper entry with its transpose's value (the pattern is
// symmetric): the factorization then sees the symmetric matrix of A's lower triangle.
void mirror_lower(int N, const int *rp, const int *ci, double *v) {
    for (int i = 0; i < N; i++)
        for (int e = rp[i]; e < rp[i + 1]; e++) {
            const int j = ci[e];
            if (j <= i) continue;
            const int *b = ci + rp[j], *end = ci + rp[j + 1];
            const int *pos = std::is_sorted(b, end) ? std::lower_bound(b, end, i) : std::find(b, end, i);
            if (pos != end && *pos == i) v[e] = v[rp[j] + (pos - b)];
        }
}

// For IC: the U position of every L entry's transpose.  L entry e = the t-th lower entry
// of row j (column i) in pattern order; U row i holds its inverted diagonal first, then
// its upper entries in pattern order.  Fails (false) if some (i, j) is missing.
bool ic_upper_map(int N, const int *fr, const int *fc, const std::vector<int> &lrp, const std::vector<int> &urp,
                  std::vector<int> &umap, std::vector<int> &lcol) {
    std::vector<std::vector<std::pair<int, int>>> up((size_t)N);
    for (int i = 0; i < N; i++) {
        int t = 0;
        for (int a = fr[i]; a < fr[i + 1]; a++)
            if (fc[a] > i) up[i].push_back({fc[a], urp[i] + 1 + t++});
        std::sort(up[i].begin(), up[i].end());
    }
    umap.assign((size_t)lrp[N], -1);
    lcol.assign((size_t)lrp[N], -1);
    for (int j = 0; j < N; j++) {
        int e = lrp[j];
        for (int a = fr[j]; a < fr[j + 1]; a++) {
            const int i = fc[a];
            if (i >= j) continue;
            auto it = std::lower_bound(up[i].begin(), up[i].end(), std::make_pair(j, -1));
            if (it == up[i].end() || it->first != j) return false;
            umap[(size_t)e] = it->second;
            lcol[(size_t)e] = i;
            e++;
        }
    }
    return true;
}

bool dilu_pattern(int N, const int *rp, const int *ci) {
    std::vector<int> stamp((size_t)N, -1);
    for (int i = 0; i < N; i++) {
        for (int a = rp[i]; a < rp[i + 1]; a++) stamp[ci[a]] = i;
        for (int a = rp[i]; a < rp[i + 1]; a++) {
            const int k = ci[a];
            if (k >= i) continue;
            for (int b = rp[k]; b < rp[k + 1]; b++) {
                const int j = ci[b];
                if (j > k && j != i && stamp[j] == i) return false;
            }
        }
    }
    return true;
}
Schedule colors(const int *cptr, int cnum, int bsize) {
    Schedule s;
    s.bsize = bsize;
    s.nlev = cnum;
    s.lev_ptr.assign(cptr, cptr + cnum + 1);
    return s;
}
// A schedule of blocks of s.bsize rows (blk: its block list, empty for the identity) as
// one of sub-blocks of g rows (g divides s.bsize): level l becomes the s.bsize / g levels
// (l, 0), (l, 1), ..., level (l, q) holding sub-block q of every block of level l.  A row
// depends on earlier levels and on rows of its own block; those in earlier sub-blocks are
// now in earlier levels, those in its own sub-block stay in-block, so this is a schedule
// of the same factor (an upper solve runs the levels backwards, sub-blocks last to first).
Schedule split_schedule(const Schedule &s, const std::vector<int> &blk, int g, std::vector<int> &sub_blk) {
    const int k = s.bsize / g;
    Schedule o;
    o.bsize = g;
    o.nlev = s.nlev * k;
    o.lev_ptr.assign((size_t)o.nlev + 1, 0);
    sub_blk.clear();
    for (int l = 0; l < s.nlev; l++)
        for (int q = 0; q < k; q++) {
            for (int i = s.lev_ptr[l]; i < s.lev_ptr[l + 1]; i++) sub_blk.push_back((blk.empty() ? i : blk[i]) * k + q);
            o.lev_ptr[(size_t)l * k + q + 1] = (int)sub_blk.size();
        }
    return o;
}
template <typename T>
int dalloc(T **p, size_t n) {
    SENK_HIP(hipMalloc((void **)p, sizeof(T) * (n ? n : 1)));
    return SENK_OK;
}
template <typename T>
int upload(senk_ctx ctx, T **p, const T *h, size_t n) {
    SENK_TRY(dalloc(p, n));
    if (n) SENK_HIP(hipMemcpyAsync(*p, h, sizeof(T) * n, hipMemcpyHostToDevice, ctx->c.stream));
    return SENK_OK;
}

// Levels of bnl-row groups of a BCSR triangular factor (for the sequential BCSR solve).
void bcsr_levels(const std::vector<int> &brptr, const std::vector<int> &bcind, int ngroups, int bnl, int bnw,
                 bool upper, std::vector<int> &lev_ptr, std::vector<int> &blk) {
    std::vector<int> depth((size_t)ngroups, 0);
    int maxd = 0;
    auto visit = [&](int g) {
        int d = 0;
        for (int b = brptr[g]; b < brptr[g + 1]; b++)
            for (int l = 0; l < bnw; l++) {
                const int og = (bcind[b] * bnw + l) / bnl;
                if ((!upper && og < g) || (upper && og > g)) d = std::max(d, depth[og] + 1);
            }
        depth[g] = d;
        maxd = std::max(maxd, d);
    };
    if (!upper) for (int g = 0; g < ngroups; g++) visit(g);
    else for (int g = ngroups - 1; g >= 0; g--) visit(g);
    if (upper) for (int g = 0; g < ngroups; g++) depth[g] = maxd - depth[g];
    const int nlev = ngroups ? maxd + 1 : 0;
    lev_ptr.assign((size_t)nlev + 1, 0);
    for (int g = 0; g < ngroups; g++) lev_ptr[depth[g] + 1]++;
    for (int l = 0; l < nlev; l++) lev_ptr[l + 1] += lev_ptr[l];
    std::vector<int> pos(lev_ptr.begin(), lev_ptr.end() - 1);
    blk.assign((size_t)ngroups, 0);
    for (int g = 0; g < ngroups; g++) blk[pos[depth[g]]++] = g;
}
}  // namespace
}  // namespace senk_amd

using namespace senk_amd;

#define CTX_CHECK(ctx) \
    do {               \
        if (!(ctx)) return SENK_ERR_ARG; \
    } while (0)

extern "C" {

const char *senk_status_string(int s) {
    switch (s) {
        case SENK_OK: return "ok";
        case SENK_ERR_ARG: return "invalid argument";
        case SENK_ERR_HIP: return "HIP runtime error";
        case SENK_ERR_ZERO_PIVOT: return "zero pivot in Ilu0";
        case SENK_ERR_DIVISIBILITY: return "size not divisible by the block size";
        case SENK_ERR_BREAKDOWN: return "Krylov breakdown (non-finite scalar)";
        case SENK_ERR_NOT_CONVERGED: return "not converged";
        case SENK_ERR_NO_DEVICE: return "no HIP device";
        case SENK_ERR_ALLOC: return "allocation failure";
        case SENK_ERR_UNSUPPORTED: return "unsupported shape";
        case SENK_ERR_SCHEDULE: return "factor entry couples two blocks of one launch";
        case SENK_ERR_TIMEOUT: return "triangular solve timed out waiting on an earlier level";
        case SENK_ERR_IO: return "file could not be read";
        case SENK_ERR_FORMAT: return "file is not a supported MatrixMarket matrix";
        default: return "unknown status";
    }
}
const char *senk_last_error(void) { return g_last_error.c_str(); }
int senk_version(void) { return 100; }

int senk_device_count(int *count) {
    if (!count) return SENK_ERR_ARG;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return SENK_OK;
}

int senk_ctx_create(int device, senk_ctx *out) {
    if (!out) return SENK_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return SENK_ERR_NO_DEVICE;
    if (device < 0 || device >= n) return SENK_ERR_ARG;
    SENK_HIP(hipSetDevice(device));
    auto *ctx = new senk_ctx_s();
    Ctx &c = ctx->c;
    c.device = device;
    SENK_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    SENK_HIP(hipMalloc(&c.red_partials, sizeof(double) * kRedSlots * kRedMaxK * kRedMaxBlocks));
    SENK_HIP(hipMalloc(&c.dscalars, sizeof(double) * 64));
    SENK_HIP(hipHostMalloc((void **)&c.hpinned, sizeof(double) * 64, hipHostMallocDefault));
    SENK_HIP(hipHostMalloc((void **)&c.hpinned_i, sizeof(int) * 64, hipHostMallocDefault));
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c.num_cus = prop.multiProcessorCount;
    // the no-allocation entry point's scratch array on this device (once per device); a
    // failure here leaves every other entry point working and makes that one report it
    (void)noalloc_init(device);
    SENK_HIP(hipDeviceSynchronize());
    *out = ctx;
    return SENK_OK;
}
int senk_ctx_destroy(senk_ctx ctx) {
    CTX_CHECK(ctx);
    Ctx &c = ctx->c;
    (void)hipSetDevice(c.device);
    (void)hipStreamSynchronize(c.stream);
    if (c.work.p) (void)hipFree(c.work.p);
    if (c.outer.p) (void)hipFree(c.outer.p);
    for (hipEvent_t e : c.ev)
        if (e) (void)hipEventDestroy(e);
    if (c.dring_h) (void)hipHostFree(c.dring_h);
    (void)hipFree(c.red_partials);
    (void)hipFree(c.dscalars);
    (void)hipHostFree(c.hpinned);
    (void)hipHostFree(c.hpinned_i);
    (void)hipStreamDestroy(c.stream);
    delete ctx;
    return SENK_OK;
}
int senk_ctx_synchronize(senk_ctx ctx) {
    CTX_CHECK(ctx);
    SENK_HIP(hipStreamSynchronize(ctx->c.stream));
    return SENK_OK;
}
int senk_device_malloc(senk_ctx ctx, size_t bytes, void **dptr) {
    CTX_CHECK(ctx);
    if (!dptr) return SENK_ERR_ARG;
    SENK_HIP(hipMalloc(dptr, bytes ? bytes : 1));
    return SENK_OK;
}
int senk_device_free(senk_ctx ctx, void *dptr) {
    CTX_CHECK(ctx);
    if (dptr) {
        SENK_HIP(hipStreamSynchronize(ctx->c.stream));
        SENK_HIP(hipFree(dptr));
        // a later allocation at this address may hold another row-pointer array
        for (auto it = ctx->nnz_cache.begin(); it != ctx->nnz_cache.end();)
            it = it->first.first == dptr ? ctx->nnz_cache.erase(it) : std::next(it);
    }
    return SENK_OK;
}
int senk_memcpy_htod(senk_ctx ctx, void *dst, const void *src, size_t bytes) {
    CTX_CHECK(ctx);
    if (!bytes) return SENK_OK;
    SENK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->c.stream));
    SENK_HIP(hipStreamSynchronize(ctx->c.stream));
    return SENK_OK;
}
int senk_memcpy_dtoh(senk_ctx ctx, void *dst, const void *src, size_t bytes) {
    CTX_CHECK(ctx);
    if (!bytes) return SENK_OK;
    SENK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->c.stream));
    SENK_HIP(hipStreamSynchronize(ctx->c.stream));
    return SENK_OK;
}
int senk_memcpy_dtod(senk_ctx ctx, void *dst, const void *src, size_t bytes) {
    CTX_CHECK(ctx);
    if (!bytes) return SENK_OK;
    SENK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->c.stream));
    return SENK_OK;
}
int senk_memset_zero(senk_ctx ctx, void *dst, size_t bytes) {
    CTX_CHECK(ctx);
    if (!bytes) return SENK_OK;
    SENK_HIP(hipMemsetAsync(dst, 0, bytes, ctx->c.stream));
    return SENK_OK;
}
int senk_event_create(senk_ctx ctx, void **ev) {
    CTX_CHECK(ctx);
    hipEvent_t e;
    SENK_HIP(hipEventCreate(&e));
    *ev = (void *)e;
    return SENK_OK;
}
int senk_event_destroy(void *ev) {
    if (ev) SENK_HIP(hipEventDestroy((hipEvent_t)ev));
    return SENK_OK;
}
int senk_event_record(senk_ctx ctx, void *ev) {
    CTX_CHECK(ctx);
    SENK_HIP(hipEventRecord((hipEvent_t)ev, ctx->c.stream));
    return SENK_OK;
}
int senk_event_elapsed_ms(void *start, void *stop, float *ms) {
    SENK_HIP(hipEventSynchronize((hipEvent_t)stop));
    SENK_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return SENK_OK;
}

// ---------------------------------------------------------------- SpMV
int senk_spmv_csr_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr, const double *dx,
                      double *dy, int N) {
    CTX_CHECK(ctx);
    if (N < 0 || (N > 0 && (!dval || !dcind || !drptr || !dx || !dy))) return SENK_ERR_ARG;
    int nnz = 0;
    if (N > 0) SENK_TRY(nnz_of(ctx, drptr, N, &nnz));
    return launch_spmv_csr<double>(&ctx->c, N, drptr, dcind, dval, dx, dy, avg_row(nnz, N));
}
int senk_spmv_csr_f32(senk_ctx ctx, const float *dval, const int *dcind, const int *drptr, const float *dx,
                      float *dy, int N) {
    CTX_CHECK(ctx);
    if (N < 0 || (N > 0 && (!dval || !dcind || !drptr || !dx || !dy))) return SENK_ERR_ARG;
    int nnz = 0;
    if (N > 0) SENK_TRY(nnz_of(ctx, drptr, N, &nnz));
    return launch_spmv_csr<float>(&ctx->c, N, drptr, dcind, dval, dx, dy, avg_row(nnz, N));
}
int senk_spmv_csr_diag_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr,
                           const double *ddiag, const double *dx, double *dy, int N) {
    CTX_CHECK(ctx);
    if (N < 0) return SENK_ERR_ARG;
    return launch_spmv_csr_diag_f64(&ctx->c, N, drptr, dcind, dval, ddiag, dx, dy);
}
int senk_spmv_bcsr_f64(senk_ctx ctx, const double *dbval, const int *dbcind, const int *dbrptr, const double *dx,
                       double *dy, int N, int bnl, int bnw) {
    CTX_CHECK(ctx);
    if (bnl <= 0 || bnw <= 0 || N < 0) return SENK_ERR_ARG;
    if (N % bnl || N % bnw) return SENK_ERR_DIVISIBILITY;
    return launch_spmv_bcsr<double>(&ctx->c, N, bnl, bnw, dbrptr, dbcind, dbval, dx, dy);
}
int senk_spmv_bcsr_f32(senk_ctx ctx, const float *dbval, const int *dbcind, const int *dbrptr, const float *dx,
                       float *dy, int N, int bnl, int bnw) {
    CTX_CHECK(ctx);
    if (bnl <= 0 || bnw <= 0 || N < 0) return SENK_ERR_ARG;
    if (N % bnl || N % bnw) return SENK_ERR_DIVISIBILITY;
    return launch_spmv_bcsr<float>(&ctx->c, N, bnl, bnw, dbrptr, dbcind, dbval, dx, dy);
}
int senk_spmv_sell_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *dwid, int len,
                       const double *dx, double *dy, int N) {
    CTX_CHECK(ctx);
    if (len <= 0 || N < 0) return SENK_ERR_ARG;
    const int ns = (N + len - 1) / len;
    return launch_spmv_sell<double>(&ctx->c, N, len, ns, dwid, dcind, dval, nullptr, dx, dy, N % len != 0);
}
int senk_spmv_sell_f32(senk_ctx ctx, const float *dval, const int *dcind, const int *dwid, int len,
                       const float *dx, float *dy, int N) {
    CTX_CHECK(ctx);
    if (len <= 0 || N < 0) return SENK_ERR_ARG;
    const int ns = (N + len - 1) / len;
    return launch_spmv_sell<float>(&ctx->c, N, len, ns, dwid, dcind, dval, nullptr, dx, dy, N % len != 0);
}
int senk_sellcs_upload(senk_ctx ctx, const senk_sellcs *h, int want_f32, senk_sellcs *d) {
    CTX_CHECK(ctx);
    if (!h || !d) return SENK_ERR_ARG;
    std::memset(d, 0, sizeof(*d));
    d->N = h->N;
    d->C = h->C;
    d->sigma = h->sigma;
    d->nslices = h->nslices;
    d->nnz_padded = h->nnz_padded;
    SENK_TRY(upload(ctx, &d->slice_ptr, h->slice_ptr, (size_t)h->nslices + 1));
    SENK_TRY(upload(ctx, &d->cind, h->cind, (size_t)h->nnz_padded));
    SENK_TRY(upload(ctx, &d->perm, h->perm, (size_t)h->nslices * h->C));
    if (h->val64) SENK_TRY(upload(ctx, &d->val64, h->val64, (size_t)h->nnz_padded));
    if (want_f32 && d->val64) {
        SENK_TRY(dalloc(&d->val32, (size_t)h->nnz_padded));
        SENK_TRY(launch_convert<double, float>(&ctx->c, h->nnz_padded, d->val64, d->val32, nullptr));
    }
    SENK_HIP(hipStreamSynchronize(ctx->c.stream));
    return SENK_OK;
}
int senk_sellcs_free_device(senk_ctx ctx, senk_sellcs *d) {
    CTX_CHECK(ctx);
    if (!d) return SENK_ERR_ARG;
    SENK_HIP(hipStreamSynchronize(ctx->c.stream));
    (void)hipFree(d->slice_ptr);
    (void)hipFree(d->cind);
    (void)hipFree(d->perm);
    (void)hipFree(d->val64);
    (void)hipFree(d->val32);
    std::memset(d, 0, sizeof(*d));
    return SENK_OK;
}
static double sell_avg_width(const senk_sellcs &d) {
    return d.nslices > 0 ? (double)d.nnz_padded / ((double)d.nslices * d.C) : 0.0;
}
int senk_spmv_sellcs_f64(senk_ctx ctx, const senk_sellcs *d, const double *dx, double *dy) {
    CTX_CHECK(ctx);
    if (!d || !d->val64) return SENK_ERR_ARG;
    return launch_spmv_sell<double>(&ctx->c, d->N, d->C, d->nslices, d->slice_ptr, d->cind, d->val64, d->perm, dx,
                                    dy, false, nullptr, sell_avg_width(*d), nullptr, false, d->nnz_padded);
}
int senk_spmv_sellcs_f32(senk_ctx ctx, const senk_sellcs *d, const float *dx, float *dy) {
    CTX_CHECK(ctx);
    if (!d || !d->val32) return SENK_ERR_ARG;
    return launch_spmv_sell<float>(&ctx->c, d->N, d->C, d->nslices, d->slice_ptr, d->cind, d->val32, d->perm, dx,
                                   dy, false, nullptr, sell_avg_width(*d), nullptr, false, d->nnz_padded);
}

// ---------------------------------------------------------------- SpTRSV
#define TRSV_ABMC(NAME, T, UP)                                                                                   \
    int NAME(senk_ctx ctx, const T *dval, const int *dcind, const int *drptr, const T *dx, T *dy, int N,         \
             const int *cptr, int cnum, int bsize) {                                                             \
        CTX_CHECK(ctx);                                                                                          \
        if (!cptr || cnum < 0 || bsize <= 0 || N < 0) return SENK_ERR_ARG;                                       \
        if (N % bsize) return SENK_ERR_DIVISIBILITY;                                                             \
        if (cptr[cnum] * bsize > N) return SENK_ERR_ARG;                                                         \
        Schedule s = colors(cptr, cnum, bsize);                                                                  \
        return launch_sptrsv<T>(&ctx->c, s, UP, drptr, dcind, dval, dx, dy, nullptr, nullptr);                   \
    }
TRSV_ABMC(senk_sptrsv_csr_l_abmc_f64, double, false)
TRSV_ABMC(senk_sptrsv_csr_u_abmc_f64, double, true)
TRSV_ABMC(senk_sptrsv_csr_l_abmc_f32, float, false)
TRSV_ABMC(senk_sptrsv_csr_u_abmc_f32, float, true)

int senk_sptrsv_csr_l_amc_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr, const double *dx,
                              double *dy, int N, const int *cptr, int cnum) {
    return senk_sptrsv_csr_l_abmc_f64(ctx, dval, dcind, drptr, dx, dy, N, cptr, cnum, 1);
}
int senk_sptrsv_csr_u_amc_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr, const double *dx,
                              double *dy, int N, const int *cptr, int cnum) {
    return senk_sptrsv_csr_u_abmc_f64(ctx, dval, dcind, drptr, dx, dy, N, cptr, cnum, 1);
}
int senk_sptrsv_csr_l_amc_f32(senk_ctx ctx, const float *dval, const int *dcind, const int *drptr, const float *dx,
                              float *dy, int N, const int *cptr, int cnum) {
    return senk_sptrsv_csr_l_abmc_f32(ctx, dval, dcind, drptr, dx, dy, N, cptr, cnum, 1);
}
int senk_sptrsv_csr_u_amc_f32(senk_ctx ctx, const float *dval, const int *dcind, const int *drptr, const float *dx,
                              float *dy, int N, const int *cptr, int cnum) {
    return senk_sptrsv_csr_u_abmc_f32(ctx, dval, dcind, drptr, dx, dy, N, cptr, cnum, 1);
}

int senk_trsv_plan_create(senk_ctx ctx, const int *rptr, const int *cind, int N, int upper, senk_trsv_plan *out) {
    CTX_CHECK(ctx);
    if (!rptr || !cind || !out || N < 0) return SENK_ERR_ARG;
    auto *p = new senk_trsv_plan_s();
    p->c = &ctx->c;
    std::vector<int> blk;
    p->s.bsize = 1;
    p->s.nlev = host::levels(rptr, cind, N, 1, upper != 0, p->s.lev_ptr, blk);
    const int st = upload(ctx, &p->s.dblk, blk.data(), blk.size());
    if (st != SENK_OK) {
        delete p;
        return st;
    }
    SENK_HIP(hipStreamSynchronize(ctx->c.stream));
    *out = p;
    return SENK_OK;
}
int senk_trsv_plan_levels(senk_trsv_plan plan) { return plan ? plan->s.nlev : -1; }
int senk_trsv_plan_destroy(senk_trsv_plan plan) {
    if (!plan) return SENK_ERR_ARG;
    (void)hipFree(plan->s.dblk);
    delete plan;
    return SENK_OK;
}
}  // extern "C"
template <typename T>
static int plan_trsv(senk_ctx ctx, senk_trsv_plan plan, bool upper, const T *dval, const int *dcind,
                     const int *drptr, const T *dx, T *dy, int N) {
    CTX_CHECK(ctx);
    if (!plan || N < 0) return SENK_ERR_ARG;
    return launch_sptrsv<T>(&ctx->c, plan->s, upper, drptr, dcind, dval, dx, dy, nullptr, nullptr);
}
extern "C" {
int senk_sptrsv_csr_l_f64(senk_ctx ctx, senk_trsv_plan plan, const double *dval, const int *dcind, const int *drptr,
                          const double *dx, double *dy, int N) {
    return plan_trsv<double>(ctx, plan, false, dval, dcind, drptr, dx, dy, N);
}
int senk_sptrsv_csr_u_f64(senk_ctx ctx, senk_trsv_plan plan, const double *dval, const int *dcind, const int *drptr,
                          const double *dx, double *dy, int N) {
    return plan_trsv<double>(ctx, plan, true, dval, dcind, drptr, dx, dy, N);
}
int senk_sptrsv_csr_l_f32(senk_ctx ctx, senk_trsv_plan plan, const float *dval, const int *dcind, const int *drptr,
                          const float *dx, float *dy, int N) {
    return plan_trsv<float>(ctx, plan, false, dval, dcind, drptr, dx, dy, N);
}
int senk_sptrsv_csr_u_f32(senk_ctx ctx, senk_trsv_plan plan, const float *dval, const int *dcind, const int *drptr,
                          const float *dx, float *dy, int N) {
    return plan_trsv<float>(ctx, plan, true, dval, dcind, drptr, dx, dy, N);
}
int senk_sptrsv_csr_l_bj_f64(senk_ctx ctx, senk_trsv_plan plan, const double *dval, const int *dcind,
                             const int *drptr, const double *dx, double *dy, int N, int bnum) {
    if (bnum <= 0) return SENK_ERR_ARG;
    return plan_trsv<double>(ctx, plan, false, dval, dcind, drptr, dx, dy, N);
}
int senk_sptrsv_csr_u_bj_f64(senk_ctx ctx, senk_trsv_plan plan, const double *dval, const int *dcind,
                             const int *drptr, const double *dx, double *dy, int N, int bnum) {
    if (bnum <= 0) return SENK_ERR_ARG;
    return plan_trsv<double>(ctx, plan, true, dval, dcind, drptr, dx, dy, N);
}
int senk_sptrsv_csr_l_bj_f32(senk_ctx ctx, senk_trsv_plan plan, const float *dval, const int *dcind,
                             const int *drptr, const float *dx, float *dy, int N, int bnum) {
    if (bnum <= 0) return SENK_ERR_ARG;
    return plan_trsv<float>(ctx, plan, false, dval, dcind, drptr, dx, dy, N);
}
int senk_sptrsv_csr_u_bj_f32(senk_ctx ctx, senk_trsv_plan plan, const float *dval, const int *dcind,
                             const int *drptr, const float *dx, float *dy, int N, int bnum) {
    if (bnum <= 0) return SENK_ERR_ARG;
    return plan_trsv<float>(ctx, plan, true, dval, dcind, drptr, dx, dy, N);
}
}  // extern "C"
template <typename T>
static int bcsr_abmc(senk_ctx ctx, const T *dbval, const int *dbcind, const int *dbrptr, const T *dx, T *dy, int N,
                     const int *cptr, int cnum, int bsize, int bnl, int bnw, bool upper) {
    CTX_CHECK(ctx);
    if (!cptr || cnum < 0 || bsize <= 0 || bnl <= 0 || bnw <= 0) return SENK_ERR_ARG;
    if (N % bsize || bsize % bnl || bnl % bnw) return SENK_ERR_DIVISIBILITY;
    Schedule s = colors(cptr, cnum, bsize);
    return launch_sptrsv_bcsr<T>(&ctx->c, s, upper, bnl, bnw, dbrptr, dbcind, dbval, dx, dy);
}
// Sequential BCSR forms: the group-level dependency levels are computed from a host
// copy of the block pattern on every call (the reference runs these serially too).
static int bcsr_seq_schedule(senk_ctx ctx, const int *dbcind, const int *dbrptr, int N, int bnl, int bnw, bool upper,
                             Schedule *s) {
    const int ng = N / bnl;
    std::vector<int> brptr((size_t)ng + 1);
    SENK_HIP(hipMemcpyAsync(brptr.data(), dbrptr, sizeof(int) * (ng + 1), hipMemcpyDeviceToHost, ctx->c.stream));
    SENK_HIP(hipStreamSynchronize(ctx->c.stream));
    std::vector<int> bcind((size_t)std::max(brptr[ng], 1));
    if (brptr[ng])
        SENK_HIP(hipMemcpyAsync(bcind.data(), dbcind, sizeof(int) * brptr[ng], hipMemcpyDeviceToHost, ctx->c.stream));
    SENK_HIP(hipStreamSynchronize(ctx->c.stream));
    s->bsize = bnl;
    std::vector<int> blk;
    bcsr_levels(brptr, bcind, ng, bnl, bnw, upper, s->lev_ptr, blk);
    s->nlev = (int)s->lev_ptr.size() - 1;
    return upload(ctx, &s->dblk, blk.data(), blk.size());
}
template <typename T>
static int bcsr_seq(senk_ctx ctx, const T *dbval, const int *dbcind, const int *dbrptr, const T *dx, T *dy, int N,
                    int bnl, int bnw, bool upper) {
    CTX_CHECK(ctx);
    if (bnl <= 0 || bnw <= 0 || N < 0) return SENK_ERR_ARG;
    if (N % bnl || bnl % bnw) return SENK_ERR_DIVISIBILITY;
    Schedule s;
    SENK_TRY(bcsr_seq_schedule(ctx, dbcind, dbrptr, N, bnl, bnw, upper, &s));
    const int st = launch_sptrsv_bcsr<T>(&ctx->c, s, upper, bnl, bnw, dbrptr, dbcind, dbval, dx, dy);
    SENK_HIP(hipStreamSynchronize(ctx->c.stream));
    (void)hipFree(s.dblk);
    return st;
}
extern "C" {
int senk_sptrsv_bcsr_l_abmc_f64(senk_ctx ctx, const double *dbval, const int *dbcind, const int *dbrptr,
                                const double *dx, double *dy, int N, const int *cptr, int cnum, int bsize, int bnl,
                                int bnw) {
    return bcsr_abmc<double>(ctx, dbval, dbcind, dbrptr, dx, dy, N, cptr, cnum, bsize, bnl, bnw, false);
}
int senk_sptrsv_bcsr_u_abmc_f64(senk_ctx ctx, const double *dbval, const int *dbcind, const int *dbrptr,
                                const double *dx, double *dy, int N, const int *cptr, int cnum, int bsize, int bnl,
                                int bnw) {
    return bcsr_abmc<double>(ctx, dbval, dbcind, dbrptr, dx, dy, N, cptr, cnum, bsize, bnl, bnw, true);
}
int senk_sptrsv_bcsr_l_abmc_f32(senk_ctx ctx, const float *dbval, const int *dbcind, const int *dbrptr,
                                const float *dx, float *dy, int N, const int *cptr, int cnum, int bsize, int bnl,
                                int bnw) {
    return bcsr_abmc<float>(ctx, dbval, dbcind, dbrptr, dx, dy, N, cptr, cnum, bsize, bnl, bnw, false);
}
int senk_sptrsv_bcsr_u_abmc_f32(senk_ctx ctx, const float *dbval, const int *dbcind, const int *dbrptr,
                                const float *dx, float *dy, int N, const int *cptr, int cnum, int bsize, int bnl,
                                int bnw) {
    return bcsr_abmc<float>(ctx, dbval, dbcind, dbrptr, dx, dy, N, cptr, cnum, bsize, bnl, bnw, true);
}
int senk_sptrsv_bcsr_l_f64(senk_ctx ctx, const double *dbval, const int *dbcind, const int *dbrptr, const double *dx,
                           double *dy, int N, int bnl, int bnw) {
    return bcsr_seq<double>(ctx, dbval, dbcind, dbrptr, dx, dy, N, bnl, bnw, false);
}
int senk_sptrsv_bcsr_u_f64(senk_ctx ctx, const double *dbval, const int *dbcind, const int *dbrptr, const double *dx,
                           double *dy, int N, int bnl, int bnw) {
    return bcsr_seq<double>(ctx, dbval, dbcind, dbrptr, dx, dy, N, bnl, bnw, true);
}
int senk_sptrsv_bcsr_l_f32(senk_ctx ctx, const float *dbval, const int *dbcind, const int *dbrptr, const float *dx,
                           float *dy, int N, int bnl, int bnw) {
    return bcsr_seq<float>(ctx, dbval, dbcind, dbrptr, dx, dy, N, bnl, bnw, false);
}
int senk_sptrsv_bcsr_u_f32(senk_ctx ctx, const float *dbval, const int *dbcind, const int *dbrptr, const float *dx,
                           float *dy, int N, int bnl, int bnw) {
    return bcsr_seq<float>(ctx, dbval, dbcind, dbrptr, dx, dy, N, bnl, bnw, true);
}

// ---------------------------------------------------------------- BLAS-1
int senk_copy_f64(senk_ctx ctx, const double *dx, double *dy, int N) {
    CTX_CHECK(ctx);
    return launch_copy<double>(&ctx->c, N, dx, dy);
}
int senk_scal_f64(senk_ctx ctx, double a, double *dx, int N) {
    CTX_CHECK(ctx);
    return launch_scal<double>(&ctx->c, N, a, dx);
}
int senk_axpy_f64(senk_ctx ctx, double a, const double *dx, double *dy, int N) {
    CTX_CHECK(ctx);
    return launch_axpby<double>(&ctx->c, N, a, dx, 1.0, dy);
}
int senk_axpby_f64(senk_ctx ctx, double a, const double *dx, double b, double *dy, int N) {
    CTX_CHECK(ctx);
    // b == 1 runs y += a*x, which is a*x + 1*y bit for bit
    return launch_axpby<double>(&ctx->c, N, a, dx, b, dy);
}
int senk_axpyz_f64(senk_ctx ctx, double a, const double *dx, const double *dy, double *dz, int N) {
    CTX_CHECK(ctx);
    return launch_axpyz<double>(&ctx->c, N, a, dx, dy, dz);
}
int senk_dot_f64_dev(senk_ctx ctx, const double *dx, const double *dy, int N, double *dres) {
    CTX_CHECK(ctx);
    if (N <= 0) {
        SENK_HIP(hipMemsetAsync(dres, 0, sizeof(double), ctx->c.stream));
        return SENK_OK;
    }
    return launch_multidot<double>(&ctx->c, N, 1, dx, (size_t)N, dy, dres);
}
int senk_dot_f64(senk_ctx ctx, const double *dx, const double *dy, int N, double *result) {
    CTX_CHECK(ctx);
    if (!result) return SENK_ERR_ARG;
    SENK_TRY(senk_dot_f64_dev(ctx, dx, dy, N, ctx->c.dscalars));
    SENK_HIP(hipMemcpyAsync(ctx->c.hpinned, ctx->c.dscalars, sizeof(double), hipMemcpyDeviceToHost, ctx->c.stream));
    SENK_HIP(hipStreamSynchronize(ctx->c.stream));
    *result = ctx->c.hpinned[0];
    return SENK_OK;
}
int senk_nrm2_f64_dev(senk_ctx ctx, const double *dx, int N, double *dres) { return senk_dot_f64_dev(ctx, dx, dx, N, dres); }
int senk_nrm2_f64(senk_ctx ctx, const double *dx, int N, double *result) {
    SENK_TRY(senk_dot_f64(ctx, dx, dx, N, result));
    *result = std::sqrt(*result);
    return SENK_OK;
}
int senk_hadprod_f64(senk_ctx ctx, const double *dx, double *dy, int N) {
    CTX_CHECK(ctx);
    return launch_hadprod<double>(&ctx->c, N, dx, dy);
}
int senk_haddiv_f64(senk_ctx ctx, const double *dx, double *dy, int N) {
    CTX_CHECK(ctx);
    return launch_haddiv<double>(&ctx->c, N, dx, dy);
}
int senk_copy_f32(senk_ctx ctx, const float *dx, float *dy, int N) {
    CTX_CHECK(ctx);
    return launch_copy<float>(&ctx->c, N, dx, dy);
}
int senk_scal_f32(senk_ctx ctx, float a, float *dx, int N) {
    CTX_CHECK(ctx);
    return launch_scal<float>(&ctx->c, N, a, dx);
}
int senk_axpy_f32(senk_ctx ctx, float a, const float *dx, float *dy, int N) {
    CTX_CHECK(ctx);
    return launch_axpby<float>(&ctx->c, N, a, dx, 1.0, dy);
}
int senk_axpby_f32(senk_ctx ctx, float a, const float *dx, float b, float *dy, int N) {
    CTX_CHECK(ctx);
    return launch_axpby<float>(&ctx->c, N, a, dx, b, dy);
}
int senk_axpyz_f32(senk_ctx ctx, float a, const float *dx, const float *dy, float *dz, int N) {
    CTX_CHECK(ctx);
    return launch_axpyz<float>(&ctx->c, N, a, dx, dy, dz);
}
int senk_hadprod_f32(senk_ctx ctx, const float *dx, float *dy, int N) {
    CTX_CHECK(ctx);
    return launch_hadprod<float>(&ctx->c, N, dx, dy);
}
int senk_haddiv_f32(senk_ctx ctx, const float *dx, float *dy, int N) {
    CTX_CHECK(ctx);
    return launch_haddiv<float>(&ctx->c, N, dx, dy);
}
int senk_dot_f32(senk_ctx ctx, const float *dx, const float *dy, int N, double *result) {
    CTX_CHECK(ctx);
    if (!result) return SENK_ERR_ARG;
    if (N <= 0) {
        *result = 0;
        return SENK_OK;
    }
    SENK_TRY(launch_multidot<float>(&ctx->c, N, 1, dx, (size_t)N, dy, ctx->c.dscalars));
    SENK_HIP(hipMemcpyAsync(ctx->c.hpinned, ctx->c.dscalars, sizeof(double), hipMemcpyDeviceToHost, ctx->c.stream));
    SENK_HIP(hipStreamSynchronize(ctx->c.stream));
    *result = ctx->c.hpinned[0];
    return SENK_OK;
}
int senk_nrm2_f32(senk_ctx ctx, const float *dx, int N, double *result) {
    SENK_TRY(senk_dot_f32(ctx, dx, dx, N, result));
    *result = std::sqrt(*result);
    return SENK_OK;
}
int senk_convert_f64_f32(senk_ctx ctx, const double *din, float *dout, int n) {
    CTX_CHECK(ctx);
    return launch_convert<double, float>(&ctx->c, n, din, dout, nullptr);
}
int senk_convert_f32_f64(senk_ctx ctx, const float *din, double *dout, int n) {
    CTX_CHECK(ctx);
    return launch_convert<float, double>(&ctx->c, n, din, dout, nullptr);
}
int senk_scaling_dev(senk_ctx ctx, double *dval, const int *dcind, const int *drptr, const double *dld,
                     const double *drd, int N) {
    CTX_CHECK(ctx);
    return launch_scaling_f64(&ctx->c, N, drptr, dcind, dval, dld, drd);
}

// ---------------------------------------------------------------- ILU(0) numeric
}  // extern "C"
template <typename T>
static int ilu0_impl(senk_ctx ctx, T *dval, const int *dcind, const int *drptr, int N, const int *cptr, int cnum,
                     int bsize, const int *hrptr, const int *hcind) {
    CTX_CHECK(ctx);
    if (N < 0 || !dval || !dcind || !drptr) return SENK_ERR_ARG;
    Ctx *c = &ctx->c;
    Schedule s;
    std::vector<int> blk;
    if (cptr && cnum > 0) {
        if (bsize <= 0) return SENK_ERR_ARG;
        if (N % bsize) return SENK_ERR_DIVISIBILITY;
        s = colors(cptr, cnum, bsize);
        if (hrptr && hcind && !host::schedule_ok(hrptr, hcind, N, s, blk, true, false)) return SENK_ERR_SCHEDULE;
    } else {
        if (!hrptr || !hcind) return SENK_ERR_ARG;
        s.bsize = 1;
        s.nlev = host::levels(hrptr, hcind, N, 1, false, s.lev_ptr, blk);
        SENK_TRY(upload(ctx, &s.dblk, blk.data(), blk.size()));
    }
    int *dpos = nullptr, *derr = nullptr;
    SENK_TRY(dalloc(&dpos, (size_t)N));
    SENK_TRY(dalloc(&derr, 1));
    SENK_HIP(hipMemsetAsync(derr, 0, sizeof(int), c->stream));
    SENK_TRY(launch_find_diag(c, N, drptr, dcind, dpos));
    SENK_TRY(launch_ilu0<T>(c, s, N, drptr, dcind, dpos, dval, derr));
    int herr = 0;
    SENK_HIP(hipMemcpyAsync(&herr, derr, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    SENK_HIP(hipStreamSynchronize(c->stream));
    (void)hipFree(dpos);
    (void)hipFree(derr);
    if (s.dblk) (void)hipFree(s.dblk);
    if (herr == 1) return SENK_ERR_ZERO_PIVOT;
    if (herr == 2) return SENK_ERR_UNSUPPORTED;
    return SENK_OK;
}
extern "C" {
int senk_ilu0_f64(senk_ctx ctx, double *dval, const int *dcind, const int *drptr, int N, const int *cptr, int cnum,
                  int bsize, const int *hrptr, const int *hcind) {
    return ilu0_impl<double>(ctx, dval, dcind, drptr, N, cptr, cnum, bsize, hrptr, hcind);
}
int senk_ilu0_f32(senk_ctx ctx, float *dval, const int *dcind, const int *drptr, int N, const int *cptr, int cnum,
                  int bsize, const int *hrptr, const int *hcind) {
    return ilu0_impl<float>(ctx, dval, dcind, drptr, N, cptr, cnum, bsize, hrptr, hcind);
}

// ---------------------------------------------------------------- solver twins
static int raw_mat(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr, int N, MatView *A);
int senk_relative_residual_error(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr,
                                 const double *db, const double *dx, int N, int print, double *relres) {
    CTX_CHECK(ctx);
    if (N <= 0 || !dval || !dcind || !drptr || !db || !dx || !relres) return SENK_ERR_ARG;
    double nb = 0;
    SENK_TRY(senk_nrm2_f64(ctx, db, N, &nb));
    MatView A;
    SENK_TRY(raw_mat(ctx, dval, dcind, drptr, N, &A));
    SENK_TRY(true_relres(&ctx->c, A, db, dx, nb, relres));
    if (print) {
        std::printf("# test %e\n", *relres);
        std::fflush(stdout);
    }
    return SENK_OK;
}
int senk_relative_residual_error_noalloc(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr,
                                         const double *db, const double *dx, int N, double *drelres) {
    CTX_CHECK(ctx);
    if (N <= 0 || !dval || !dcind || !drptr || !db || !dx || !drelres) return SENK_ERR_ARG;
    return relres_noalloc(&ctx->c, dval, dcind, drptr, db, dx, N, drelres);
}
static int finish(senk_ctx ctx, const MatView &A, const double *db, const double *dx, double nrm_b, senk_result *res) {
    double tr = 0;
    SENK_TRY(true_relres(&ctx->c, A, db, dx, nrm_b, &tr));
    res->true_relres = tr;
    return res->status;
}
static int raw_mat(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr, int N, MatView *A) {
    A->N = N;
    A->rptr = drptr;
    A->cind = dcind;
    A->v64 = dval;
    int nnz = 0;
    SENK_TRY(nnz_of(ctx, drptr, N, &nnz));
    A->nnz = nnz;
    A->avg_row = avg_row(nnz, N);
    return SENK_OK;
}
static int run_twin(senk_ctx ctx, const MatView &A, const PrecView &M, SolveParams &p, const double *db, double *dx,
                    senk_result *res, double *hist, int hist_cap) {
    senk_result tmp;
    if (!res) res = &tmp;
    SENK_TRY(krylov_solve<double>(&ctx->c, A, M, db, dx, p, res, hist, hist_cap));
    return finish(ctx, A, db, dx, p.nrm_b, res);
}

int senk_gmresm_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr, const double *db, double *dx,
                    double nrm_b, int outer, int m, int N, double epsilon, int verbose, senk_result *res, double *hist,
                    int hist_cap) {
    CTX_CHECK(ctx);
    if (N <= 0 || m <= 0 || outer < 0) return SENK_ERR_ARG;
    MatView A;
    SENK_TRY(raw_mat(ctx, dval, dcind, drptr, N, &A));
    PrecView M;
    M.N = N;
    SolveParams p;
    p.solver = SENK_GMRES;
    p.m = m;
    p.outer = outer;
    p.eps = epsilon;
    p.nrm_b = nrm_b;
    p.verbose = verbose;
    return run_twin(ctx, A, M, p, db, dx, res, hist, hist_cap);
}
static void ilu_view(PrecView &M, int N, const double *dlval, const int *dlcind, const int *dlrptr, const double *duval,
                     const int *ducind, const int *durptr) {
    M.N = N;
    M.kind = 1;
    M.l64 = dlval;
    M.lcind = dlcind;
    M.lrptr = dlrptr;
    M.u64 = duval;
    M.ucind = ducind;
    M.urptr = durptr;
}
int senk_ilu_gmresm_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr, const double *dlval,
                        const int *dlcind, const int *dlrptr, const double *duval, const int *ducind,
                        const int *durptr, senk_trsv_plan lplan, senk_trsv_plan uplan, const double *db, double *dx,
                        double nrm_b, int outer, int m, int N, double epsilon, int verbose, senk_result *res,
                        double *hist, int hist_cap) {
    CTX_CHECK(ctx);
    if (N <= 0 || m <= 0 || !lplan || !uplan) return SENK_ERR_ARG;
    MatView A;
    SENK_TRY(raw_mat(ctx, dval, dcind, drptr, N, &A));
    PrecView M;
    ilu_view(M, N, dlval, dlcind, dlrptr, duval, ducind, durptr);
    M.ls = lplan->s;
    M.us = uplan->s;
    SolveParams p;
    p.solver = SENK_GMRES;
    p.m = m;
    p.outer = outer;
    p.eps = epsilon;
    p.nrm_b = nrm_b;
    p.verbose = verbose;
    return run_twin(ctx, A, M, p, db, dx, res, hist, hist_cap);
}
int senk_abmc_ilu_gmresm_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr,
                             const double *dlval, const int *dlcind, const int *dlrptr, const double *duval,
                             const int *ducind, const int *durptr, const int *cptr, int cnum, int bsize,
                             const double *db, double *dx, double nrm_b, int outer, int m, int N, double epsilon,
                             int verbose, senk_result *res, double *hist, int hist_cap) {
    CTX_CHECK(ctx);
    if (N <= 0 || m <= 0 || !cptr || cnum <= 0 || bsize <= 0) return SENK_ERR_ARG;
    if (N % bsize) return SENK_ERR_DIVISIBILITY;
    MatView A;
    SENK_TRY(raw_mat(ctx, dval, dcind, drptr, N, &A));
    PrecView M;
    ilu_view(M, N, dlval, dlcind, dlrptr, duval, ducind, durptr);
    M.ls = colors(cptr, cnum, bsize);
    M.us = M.ls;
    SolveParams p;
    p.solver = SENK_GMRES;
    p.m = m;
    p.outer = outer;
    p.eps = epsilon;
    p.nrm_b = nrm_b;
    p.verbose = verbose;
    return run_twin(ctx, A, M, p, db, dx, res, hist, hist_cap);
}
int senk_abmc_ilub_gmresm_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr,
                              const double *dblval, const int *dblcind, const int *dblrptr, const double *dbuval,
                              const int *dbucind, const int *dburptr, const int *cptr, int cnum, int bsize,
                              const double *db, double *dx, double nrm_b, int outer, int m, int N, double epsilon,
                              int bnl, int bnw, int verbose, senk_result *res, double *hist, int hist_cap) {
    CTX_CHECK(ctx);
    if (N <= 0 || m <= 0 || !cptr || cnum <= 0 || bsize <= 0 || bnl <= 0 || bnw <= 0) return SENK_ERR_ARG;
    if (N % bsize || bsize % bnl || bnl % bnw) return SENK_ERR_DIVISIBILITY;
    MatView A;
    SENK_TRY(raw_mat(ctx, dval, dcind, drptr, N, &A));
    PrecView M;
    M.N = N;
    M.kind = 2;
    M.bnl = bnl;
    M.bnw = bnw;
    M.bl64 = dblval;
    M.blcind = dblcind;
    M.blrptr = dblrptr;
    M.bu64 = dbuval;
    M.bucind = dbucind;
    M.burptr = dburptr;
    M.ls = colors(cptr, cnum, bsize);
    M.us = M.ls;
    SolveParams p;
    p.solver = SENK_GMRES;
    p.m = m;
    p.outer = outer;
    p.eps = epsilon;
    p.nrm_b = nrm_b;
    p.verbose = verbose;
    return run_twin(ctx, A, M, p, db, dx, res, hist, hist_cap);
}
int senk_gcrm_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr, const double *db, double *dx,
                  double nrm_b, int outer, int m, int N, double epsilon, int verbose, senk_result *res, double *hist,
                  int hist_cap) {
    CTX_CHECK(ctx);
    if (N <= 0 || m <= 0) return SENK_ERR_ARG;
    MatView A;
    SENK_TRY(raw_mat(ctx, dval, dcind, drptr, N, &A));
    PrecView M;
    M.N = N;
    SolveParams p;
    p.solver = SENK_GCR;
    p.m = m;
    p.outer = outer;
    p.eps = epsilon;
    p.nrm_b = nrm_b;
    p.verbose = verbose;
    return run_twin(ctx, A, M, p, db, dx, res, hist, hist_cap);
}
int senk_abmc_ilu_gcrm_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr, const double *dlval,
                           const int *dlcind, const int *dlrptr, const double *duval, const int *ducind,
                           const int *durptr, const int *cptr, int cnum, int bsize, const double *db, double *dx,
                           double nrm_b, int outer, int m, int N, double epsilon, int verbose, senk_result *res,
                           double *hist, int hist_cap) {
    CTX_CHECK(ctx);
    if (N <= 0 || m <= 0 || !cptr || cnum <= 0 || bsize <= 0) return SENK_ERR_ARG;
    if (N % bsize) return SENK_ERR_DIVISIBILITY;
    MatView A;
    SENK_TRY(raw_mat(ctx, dval, dcind, drptr, N, &A));
    PrecView M;
    ilu_view(M, N, dlval, dlcind, dlrptr, duval, ducind, durptr);
    M.ls = colors(cptr, cnum, bsize);
    M.us = M.ls;
    SolveParams p;
    p.solver = SENK_GCR;
    p.m = m;
    p.outer = outer;
    p.eps = epsilon;
    p.nrm_b = nrm_b;
    p.verbose = verbose;
    return run_twin(ctx, A, M, p, db, dx, res, hist, hist_cap);
}
int senk_bicgstab_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr, const double *db,
                      double *dx, double nrm_b, int max_iter, int N, double epsilon, int verbose, senk_result *res,
                      double *hist, int hist_cap) {
    CTX_CHECK(ctx);
    if (N <= 0 || max_iter < 0) return SENK_ERR_ARG;
    MatView A;
    SENK_TRY(raw_mat(ctx, dval, dcind, drptr, N, &A));
    PrecView M;
    M.N = N;
    SolveParams p;
    p.solver = SENK_BICGSTAB;
    p.max_iter = max_iter;
    p.eps = epsilon;
    p.nrm_b = nrm_b;
    p.verbose = verbose;
    return run_twin(ctx, A, M, p, db, dx, res, hist, hist_cap);
}
int senk_abmc_ilu_bicgstab_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr,
                               const double *dlval, const int *dlcind, const int *dlrptr, const double *duval,
                               const int *ducind, const int *durptr, const int *cptr, int cnum, int bsize,
                               const double *db, double *dx, double nrm_b, int max_iter, int N, double epsilon,
                               int verbose, senk_result *res, double *hist, int hist_cap) {
    CTX_CHECK(ctx);
    if (N <= 0 || max_iter < 0 || !cptr || cnum <= 0 || bsize <= 0) return SENK_ERR_ARG;
    if (N % bsize) return SENK_ERR_DIVISIBILITY;
    MatView A;
    SENK_TRY(raw_mat(ctx, dval, dcind, drptr, N, &A));
    PrecView M;
    ilu_view(M, N, dlval, dlcind, dlrptr, duval, ducind, durptr);
    M.ls = colors(cptr, cnum, bsize);
    M.us = M.ls;
    SolveParams p;
    p.solver = SENK_BICGSTAB;
    p.max_iter = max_iter;
    p.eps = epsilon;
    p.nrm_b = nrm_b;
    p.verbose = verbose;
    return run_twin(ctx, A, M, p, db, dx, res, hist, hist_cap);
}

// ---------------------------------------------------------------- further driver twins
// The reference's sequential and block-Jacobi preconditioned drivers run here on level
// schedules of the factors (plans, or BCSR block-row levels computed from the pattern);
// the triangular solves give the sequential result bit for bit.
static SolveParams twin_params(int solver, int outer, int m, int max_iter, double nrm_b, double eps, int verbose) {
    SolveParams p;
    p.solver = solver;
    if (m > 0) p.m = m;  // BiCGStab has no restart length
    p.outer = outer;
    p.max_iter = max_iter;
    p.eps = eps;
    p.nrm_b = nrm_b;
    p.verbose = verbose;
    return p;
}
static int plan_twin(senk_ctx ctx, int solver, const double *dval, const int *dcind, const int *drptr,
                     const double *dlval, const int *dlcind, const int *dlrptr, const double *duval,
                     const int *ducind, const int *durptr, senk_trsv_plan lplan, senk_trsv_plan uplan,
                     const double *db, double *dx, double nrm_b, int outer, int m, int max_iter, int N,
                     double epsilon, int verbose, senk_result *res, double *hist, int hist_cap) {
    CTX_CHECK(ctx);
    if (N <= 0 || !lplan || !uplan || (solver != SENK_BICGSTAB && m <= 0)) return SENK_ERR_ARG;
    MatView A;
    SENK_TRY(raw_mat(ctx, dval, dcind, drptr, N, &A));
    PrecView M;
    ilu_view(M, N, dlval, dlcind, dlrptr, duval, ducind, durptr);
    M.ls = lplan->s;
    M.us = uplan->s;
    SolveParams p = twin_params(solver, outer, m, max_iter, nrm_b, epsilon, verbose);
    return run_twin(ctx, A, M, p, db, dx, res, hist, hist_cap);
}
static int ilub_seq_twin(senk_ctx ctx, int solver, const double *dval, const int *dcind, const int *drptr,
                         const double *dblval, const int *dblcind, const int *dblrptr, const double *dbuval,
                         const int *dbucind, const int *dburptr, const double *db, double *dx, double nrm_b,
                         int outer, int m, int max_iter, int N, double epsilon, int bnl, int bnw, int verbose,
                         senk_result *res, double *hist, int hist_cap) {
    CTX_CHECK(ctx);
    if (N <= 0 || bnl <= 0 || bnw <= 0 || (solver != SENK_BICGSTAB && m <= 0)) return SENK_ERR_ARG;
    if (N % bnl || bnl % bnw) return SENK_ERR_DIVISIBILITY;
    MatView A;
    SENK_TRY(raw_mat(ctx, dval, dcind, drptr, N, &A));
    PrecView M;
    M.N = N;
    M.kind = 2;
    M.bnl = bnl;
    M.bnw = bnw;
    M.bl64 = dblval;
    M.blcind = dblcind;
    M.blrptr = dblrptr;
    M.bu64 = dbuval;
    M.bucind = dbucind;
    M.burptr = dburptr;
    int st = bcsr_seq_schedule(ctx, dblcind, dblrptr, N, bnl, bnw, false, &M.ls);
    if (st == SENK_OK) st = bcsr_seq_schedule(ctx, dbucind, dburptr, N, bnl, bnw, true, &M.us);
    if (st == SENK_OK) {
        SolveParams p = twin_params(solver, outer, m, max_iter, nrm_b, epsilon, verbose);
        st = run_twin(ctx, A, M, p, db, dx, res, hist, hist_cap);
    }
    (void)hipStreamSynchronize(ctx->c.stream);
    (void)hipFree(M.ls.dblk);
    (void)hipFree(M.us.dblk);
    return st;
}

int senk_gmresm_bcsr_f64(senk_ctx ctx, const double *dbval, const int *dbcind, const int *dbrptr, const double *db,
                         double *dx, double nrm_b, int outer, int m, int N, double epsilon, int bnl, int bnw,
                         int verbose, senk_result *res, double *hist, int hist_cap) {
    CTX_CHECK(ctx);
    if (N <= 0 || m <= 0 || outer < 0 || bnl <= 0 || bnw <= 0 || !dbval || !dbcind || !dbrptr) return SENK_ERR_ARG;
    if (N % bnl || N % bnw) return SENK_ERR_DIVISIBILITY;
    MatView A;
    A.N = N;
    A.bnl = bnl;
    A.bnw = bnw;
    A.brptr = dbrptr;
    A.bcind = dbcind;
    A.bv64 = dbval;
    PrecView M;
    M.N = N;
    SolveParams p = twin_params(SENK_GMRES, outer, m, 0, nrm_b, epsilon, verbose);
    return run_twin(ctx, A, M, p, db, dx, res, hist, hist_cap);
}
int senk_amc_ilu_gmresm_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr,
                            const double *dlval, const int *dlcind, const int *dlrptr, const double *duval,
                            const int *ducind, const int *durptr, const int *cptr, int cnum, const double *db,
                            double *dx, double nrm_b, int outer, int m, int N, double epsilon, int verbose,
                            senk_result *res, double *hist, int hist_cap) {
    return senk_abmc_ilu_gmresm_f64(ctx, dval, dcind, drptr, dlval, dlcind, dlrptr, duval, ducind, durptr, cptr, cnum,
                                    1, db, dx, nrm_b, outer, m, N, epsilon, verbose, res, hist, hist_cap);
}
int senk_bj_ilu_gmresm_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr, const double *dlval,
                           const int *dlcind, const int *dlrptr, const double *duval, const int *ducind,
                           const int *durptr, senk_trsv_plan lplan, senk_trsv_plan uplan, int bnum, const double *db,
                           double *dx, double nrm_b, int outer, int m, int N, double epsilon, int verbose,
                           senk_result *res, double *hist, int hist_cap) {
    if (bnum <= 0) return SENK_ERR_ARG;
    return plan_twin(ctx, SENK_GMRES, dval, dcind, drptr, dlval, dlcind, dlrptr, duval, ducind, durptr, lplan, uplan,
                     db, dx, nrm_b, outer, m, 0, N, epsilon, verbose, res, hist, hist_cap);
}
int senk_ilub_gmresm_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr, const double *dblval,
                         const int *dblcind, const int *dblrptr, const double *dbuval, const int *dbucind,
                         const int *dburptr, const double *db, double *dx, double nrm_b, int outer, int m, int N,
                         double epsilon, int bnl, int bnw, int verbose, senk_result *res, double *hist, int hist_cap) {
    return ilub_seq_twin(ctx, SENK_GMRES, dval, dcind, drptr, dblval, dblcind, dblrptr, dbuval, dbucind, dburptr, db,
                         dx, nrm_b, outer, m, 0, N, epsilon, bnl, bnw, verbose, res, hist, hist_cap);
}
int senk_ilu_gcrm_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr, const double *dlval,
                      const int *dlcind, const int *dlrptr, const double *duval, const int *ducind, const int *durptr,
                      senk_trsv_plan lplan, senk_trsv_plan uplan, const double *db, double *dx, double nrm_b,
                      int outer, int m, int N, double epsilon, int verbose, senk_result *res, double *hist,
                      int hist_cap) {
    return plan_twin(ctx, SENK_GCR, dval, dcind, drptr, dlval, dlcind, dlrptr, duval, ducind, durptr, lplan, uplan,
                     db, dx, nrm_b, outer, m, 0, N, epsilon, verbose, res, hist, hist_cap);
}
int senk_ilub_gcrm_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr, const double *dblval,
                       const int *dblcind, const int *dblrptr, const double *dbuval, const int *dbucind,
                       const int *dburptr, const double *db, double *dx, double nrm_b, int outer, int m, int N,
                       double epsilon, int bnl, int bnw, int verbose, senk_result *res, double *hist, int hist_cap) {
    return ilub_seq_twin(ctx, SENK_GCR, dval, dcind, drptr, dblval, dblcind, dblrptr, dbuval, dbucind, dburptr, db, dx,
                         nrm_b, outer, m, 0, N, epsilon, bnl, bnw, verbose, res, hist, hist_cap);
}
int senk_ilu_bicgstab_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr, const double *dlval,
                          const int *dlcind, const int *dlrptr, const double *duval, const int *ducind,
                          const int *durptr, senk_trsv_plan lplan, senk_trsv_plan uplan, const double *db, double *dx,
                          double nrm_b, int max_iter, int N, double epsilon, int verbose, senk_result *res,
                          double *hist, int hist_cap) {
    if (max_iter < 0) return SENK_ERR_ARG;
    return plan_twin(ctx, SENK_BICGSTAB, dval, dcind, drptr, dlval, dlcind, dlrptr, duval, ducind, durptr, lplan,
                     uplan, db, dx, nrm_b, 0, 0, max_iter, N, epsilon, verbose, res, hist, hist_cap);
}
int senk_ilub_bicgstab_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr, const double *dblval,
                           const int *dblcind, const int *dblrptr, const double *dbuval, const int *dbucind,
                           const int *dburptr, const double *db, double *dx, double nrm_b, int max_iter, int N,
                           double epsilon, int bnl, int bnw, int verbose, senk_result *res, double *hist,
                           int hist_cap) {
    if (max_iter < 0) return SENK_ERR_ARG;
    return ilub_seq_twin(ctx, SENK_BICGSTAB, dval, dcind, drptr, dblval, dblcind, dblrptr, dbuval, dbucind, dburptr,
                         db, dx, nrm_b, 0, 0, max_iter, N, epsilon, bnl, bnw, verbose, res, hist, hist_cap);
}

// ---------------------------------------------------------------- handle API
int senk_mat_create(senk_ctx ctx, int N, const int *rptr, const int *cind, const double *val, int flags,
                    senk_mat *out) {
    CTX_CHECK(ctx);
    if (N <= 0 || !rptr || !cind || !val || !out) return SENK_ERR_ARG;
    auto *A = new senk_mat_s();
    A->ctx = ctx;
    A->N = N;
    A->nnz = rptr[N];
    int st = SENK_OK;
    do {
        if ((st = upload(ctx, &A->rptr, rptr, (size_t)N + 1))) break;
        if ((st = upload(ctx, &A->cind, cind, (size_t)A->nnz))) break;
        if ((st = upload(ctx, &A->v64, val, (size_t)A->nnz))) break;
        if (flags & 1) {
            if ((st = dalloc(&A->v32, (size_t)A->nnz))) break;
            if ((st = launch_convert<double, float>(&ctx->c, A->nnz, A->v64, A->v32, nullptr))) break;
        }
        if (flags & 2) {
            senk_sellcs h;
            if ((st = senk_csr2sellcs(val, cind, rptr, N, 64, 1024, &h))) break;
            st = senk_sellcs_upload(ctx, &h, flags & 1, &A->sell);
            senk_sellcs_free_host(&h);
            if (st) break;
            A->has_sell = true;
        }
        if (hipStreamSynchronize(ctx->c.stream) != hipSuccess) st = SENK_ERR_HIP;
    } while (0);
    if (st) {
        senk_mat_destroy(A);
        return st;
    }
    MatView &v = A->view;
    v.N = N;
    v.nnz = A->nnz;
    v.rptr = A->rptr;
    v.cind = A->cind;
    v.v64 = A->v64;
    v.v32 = A->v32;
    v.avg_row = avg_row(A->nnz, N);
    v.sell = A->has_sell ? &A->sell : nullptr;
    *out = A;
    return SENK_OK;
}
int senk_mat_destroy(senk_mat A) {
    if (!A) return SENK_ERR_ARG;
    (void)hipStreamSynchronize(A->ctx->c.stream);
    (void)hipFree(A->rptr);
    (void)hipFree(A->cind);
    (void)hipFree(A->v64);
    (void)hipFree(A->v32);
    if (A->has_sell) senk_sellcs_free_device(A->ctx, &A->sell);
    (void)hipFree(A->sell_order[0]);
    (void)hipFree(A->sell_order[1]);
    delete A;
    return SENK_OK;
}
int senk_mat_spmv_f64(senk_mat A, const double *dx, double *dy) {
    if (!A) return SENK_ERR_ARG;
    return spmv<double>(&A->ctx->c, A->view, dx, dy);
}
int senk_mat_spmv_f32(senk_mat A, const float *dx, float *dy) {
    if (!A || !A->v32) return SENK_ERR_ARG;
    return spmv<float>(&A->ctx->c, A->view, dx, dy);
}
int senk_mat_nnz(senk_mat A, int *nnz) {
    if (!A || !nnz) return SENK_ERR_ARG;
    *nnz = A->nnz;
    return SENK_OK;
}

// With a color ordering every color spans the whole domain, in the same spatial order.
// The SELL SpMV hands its slices to the workgroups so that the j-th eighth of every
// color comes in the j-th eighth of the grid, which the workgroup numbering gives to
// XCD j: an XCD then works on one region of the domain and finds the gathered entries
// of x in its own L2.  Each row is still computed alone, so results do not change.
}  // extern "C"
// Average time (ms) of `reps` SpMVs of A in one configuration (HIP events on the stream).
template <typename T>
static int time_sell(senk_mat A, const T *x, T *y, const int *order, bool nt, int reps, float *ms) {
    Ctx *c = &A->ctx->c;
    const senk_sellcs &d = A->sell;
    const T *v = std::is_same<T, double>::value ? (const T *)d.val64 : (const T *)d.val32;
    const double w = (double)d.nnz_padded / ((double)d.nslices * d.C);
    auto run = [&]() {
        return launch_spmv_sell<T>(c, d.N, d.C, d.nslices, d.slice_ptr, d.cind, v, d.perm, x, y, false, nullptr, w,
                                   order, nt, d.nnz_padded);
    };
    SENK_TRY(run());  // warm-up
    hipEvent_t e0, e1;
    SENK_HIP(hipEventCreate(&e0));
    SENK_HIP(hipEventCreate(&e1));
    int st = SENK_OK;
    if (hipEventRecord(e0, c->stream) != hipSuccess) st = SENK_ERR_HIP;
    for (int k = 0; k < reps && !st; k++) st = run();
    if (!st && (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(ms, e0, e1) != hipSuccess))
        st = SENK_ERR_HIP;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return st;
}

// The order in which the SELL SpMV hands out its slices once a colored preconditioner is
// built on A.  With an ABMC ordering every color spans the whole domain, so a contiguous
// part of the slice sequence gathers x from everywhere; here the slices are grouped by
// region instead: the k-th of R equal parts of every color (by row position inside the
// color) make region k, regions in order, colors in order inside a region.  The
// dispatcher hands workgroup b to XCD b mod 8 and each XCD works through a contiguous
// eighth of the renumbered grid (xcd_block), so an XCD covers R/8 regions of the domain
// and gathers x mostly through its own L2.  R = 8 (one region per XCD) or 64 (eight per
// XCD, each a smaller footprint of x), and plain or non-temporal loads of the matrix
// stream: which of the four is fastest depends on the matrix (measured on MI355X: 128^3
// 7-point FP64 46.4 us with 8 regions, 39.4 with 64; 27-point FP64 154.6 / 160.4 us,
// 145.9 with 8 regions and non-temporal loads), so each value type takes the fastest of
// ten products of each.  Every row keeps its sum order: the results are the same bits.
static int set_sell_order(senk_mat A, const int *cptr, int cnum, int bsize) {
    if (!A->has_sell || !cptr || cnum <= 0) return SENK_OK;
    const senk_sellcs &d = A->sell;
    const int ns = d.nslices, C = d.C;
    std::vector<int> perm((size_t)ns * C);
    Ctx *c = &A->ctx->c;
    SENK_HIP(hipMemcpyAsync(perm.data(), d.perm, perm.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    SENK_HIP(hipStreamSynchronize(c->stream));
    // position of each slice inside its color, as a fraction in [0, 1)
    std::vector<double> frac(ns);
    for (int q = 0; q < ns; q++) {
        const int row = perm[(size_t)q * C];
        int col = (int)(std::upper_bound(cptr, cptr + cnum + 1, row / bsize) - cptr) - 1;
        col = std::max(0, std::min(cnum - 1, col));
        const long long r0 = (long long)cptr[col] * bsize, r1 = (long long)cptr[col + 1] * bsize;
        frac[q] = r1 > r0 ? (double)(row - r0) / (double)(r1 - r0) : 0.0;
    }
    const int regions[2] = {8, 64};
    std::vector<int> orders[2];  // kept until the copies are done
    for (int k = 0; k < 2; k++) {
        const int R = regions[k];
        std::vector<int> reg(ns);
        std::vector<int> &order = orders[k];
        order.reserve(ns);
        for (int q = 0; q < ns; q++) reg[q] = std::max(0, std::min(R - 1, (int)(frac[q] * R)));
        for (int e = 0; e < R; e++)
            for (int q = 0; q < ns; q++)
                if (reg[q] == e) order.push_back(q);
        if (!A->sell_order[k]) SENK_TRY(dalloc(&A->sell_order[k], (size_t)ns));
        SENK_HIP(hipMemcpyAsync(A->sell_order[k], order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice,
                                c->stream));
    }
    // time the four configurations per value type on a zero vector (the matrix values
    // and the access pattern, not x, set the time)
    void *buf = nullptr;
    const size_t ld = ((size_t)A->N + 3) & ~(size_t)3;  // 16-byte aligned second vector
    SENK_TRY(ctx_work(c, 2 * ld * sizeof(double), &buf));
    double *x = static_cast<double *>(buf), *y = x + ld;
    SENK_HIP(hipMemsetAsync(x, 0, sizeof(double) * A->N, c->stream));
    for (int t = 0; t < 2; t++) {
        const bool f32 = t == 1;
        if (f32 && !d.val32) continue;
        float best = 0;
        int best_k = 0;
        bool best_nt = false;
        for (int k = 0; k < 2; k++)
            for (int nt = 0; nt < 2; nt++) {
                float ms = 0;
                SENK_TRY(f32 ? time_sell<float>(A, (const float *)x, (float *)y, A->sell_order[k], nt, 10, &ms)
                             : time_sell<double>(A, x, y, A->sell_order[k], nt, 10, &ms));
                if ((k == 0 && nt == 0) || ms < best) {
                    best = ms;
                    best_k = k;
                    best_nt = nt;
                }
            }
        if (f32) {
            A->view.sell_order32 = A->sell_order[best_k];
            A->view.sell_nt32 = best_nt;
        } else {
            A->view.sell_order64 = A->sell_order[best_k];
            A->view.sell_nt64 = best_nt;
        }
    }
    SENK_HIP(hipStreamSynchronize(c->stream));
    return SENK_OK;
}
extern "C" {

int senk_prec_create_ilu(senk_ctx ctx, senk_mat A, const int *rptr, const int *cind, const double *val,
                         const int *cptr, int cnum, int bsize, const senk_ilu_opts *opts, senk_prec *out) {
    CTX_CHECK(ctx);
    if (!A || !rptr || !cind || !val || !opts || !out) return SENK_ERR_ARG;
    const int N = A->N;
    if (bsize <= 0 || N % bsize) return SENK_ERR_DIVISIBILITY;
    Ctx *c = &ctx->c;
    // host symbolic pattern of the factor
    const int nnzA = rptr[N];
    double *fv = (double *)std::malloc(sizeof(double) * (nnzA ? nnzA : 1));
    int *fc = (int *)std::malloc(sizeof(int) * (nnzA ? nnzA : 1)), *fr = (int *)std::malloc(sizeof(int) * (N + 1));
    std::memcpy(fv, val, sizeof(double) * nnzA);
    std::memcpy(fc, cind, sizeof(int) * nnzA);
    std::memcpy(fr, rptr, sizeof(int) * (N + 1));
    int st = SENK_OK;
    const bool ic = opts->kind == SENK_IC0 || opts->kind == SENK_ICP;
    if (opts->kind < SENK_ILU0 || opts->kind > SENK_ICP) st = SENK_ERR_ARG;
    else if (ic && !symmetric_csr(N, rptr, cind, val, kSymTol)) {
        record_error("IC(0)/IC(p): the matrix is not symmetric");
        st = SENK_ERR_ARG;
    } else if ((opts->kind == SENK_ILUP || opts->kind == SENK_ICP) && opts->p > 0)
        st = senk_alloc_level_zero(&fv, &fc, &fr, N, opts->p);
    else if (opts->kind == SENK_ILUB) st = senk_alloc_block_zero(&fv, &fc, &fr, N, opts->bnl, opts->bnw);
    if (st) {
        std::free(fv); std::free(fc); std::free(fr);
        return st;
    }
    // IC: factor the symmetric matrix of A's lower triangle (A may differ from its
    // transpose in the last bits, e.g. after Jacobi scaling)
    if (ic) mirror_lower(N, fr, fc, fv);
    if ((st = set_sell_order(A, cptr, cnum, bsize))) {
        std::free(fv); std::free(fc); std::free(fr);
        return st;
    }
    auto *M = new senk_prec_s();
    M->ctx = ctx;
    PrecView &v = M->view;
    v.N = N;
    v.kind = 1;
    const int fnnz = fr[N];
    // schedules: the colors when they keep every factor entry out of the blocks of its
    // own launch, otherwise the dependency levels of the factor's blocks
    std::vector<int> noblk, lblk, ublk;
    Schedule cs;
    bool use_colors = cptr && cnum > 0;
    if (use_colors) {
        cs = colors(cptr, cnum, bsize);
        use_colors = host::schedule_ok(fr, fc, N, cs, noblk, true, true);
    }
    if (use_colors) {
        v.ls = cs;
        v.us = cs;
    } else {
        v.ls.bsize = v.us.bsize = bsize;
        v.ls.nlev = host::levels(fr, fc, N, bsize, false, v.ls.lev_ptr, lblk);
        v.us.nlev = host::levels(fr, fc, N, bsize, true, v.us.lev_ptr, ublk);
    }
    // host split pattern (row pointers)
    M->hlrptr.assign((size_t)N + 1, 0);
    M->hurptr.assign((size_t)N + 1, 0);
    for (int i = 0; i < N; i++) {
        int nl = 0, nu = 0;
        for (int j = fr[i]; j < fr[i + 1]; j++) (fc[j] < i ? nl : nu)++;
        M->hlrptr[i + 1] = M->hlrptr[i] + nl;
        M->hurptr[i + 1] = M->hurptr[i] + nu;
    }
    M->lnnz = M->hlrptr[N];
    M->unnz = M->hurptr[N];
    int *dfr = nullptr, *dfc = nullptr;
    int *dpos = nullptr;  // diagonal positions of the combined factor (kept for IC)
    double *dfv = nullptr;
    int *lr = nullptr, *lc = nullptr, *ur = nullptr, *uc = nullptr;
    double *l64 = nullptr, *u64 = nullptr;
    do {
        if ((st = upload(ctx, &dfr, fr, (size_t)N + 1))) break;
        if ((st = upload(ctx, &dfc, fc, (size_t)fnnz))) break;
        if ((st = upload(ctx, &dfv, fv, (size_t)fnnz))) break;
        if (!lblk.empty()) {
            if ((st = upload(ctx, &v.ls.dblk, lblk.data(), lblk.size()))) break;
            M->owned.push_back(v.ls.dblk);
        }
        if (!ublk.empty()) {
            if ((st = upload(ctx, &v.us.dblk, ublk.data(), ublk.size()))) break;
            M->owned.push_back(v.us.dblk);
        }
        // numeric factorization on the device (dependency-respecting schedule)
        int *derr = nullptr;
        if ((st = dalloc(&dpos, (size_t)N))) break;
        if ((st = dalloc(&derr, 1))) break;
        if (hipMemsetAsync(derr, 0, sizeof(int), c->stream) != hipSuccess) { st = SENK_ERR_HIP; break; }
        if ((st = launch_find_diag(c, N, dfr, dfc, dpos))) break;
        if ((st = launch_ilu0<double>(c, v.ls, N, dfr, dfc, dpos, dfv, derr))) break;
        int herr = 0;
        if (hipMemcpyAsync(&herr, derr, sizeof(int), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess) { st = SENK_ERR_HIP; break; }
        (void)hipFree(derr);
        if (herr == 1) { st = SENK_ERR_ZERO_PIVOT; break; }
        if (herr == 2) { st = SENK_ERR_UNSUPPORTED; break; }
        // split into L (strict lower) and U (inverted diagonal first)
        if ((st = upload(ctx, &lr, M->hlrptr.data(), (size_t)N + 1))) break;
        if ((st = upload(ctx, &ur, M->hurptr.data(), (size_t)N + 1))) break;
        if ((st = dalloc(&lc, (size_t)M->lnnz))) break;
        if ((st = dalloc(&uc, (size_t)M->unnz))) break;
        if ((st = dalloc(&l64, (size_t)M->lnnz))) break;
        if ((st = dalloc(&u64, (size_t)M->unnz))) break;
        M->owned.insert(M->owned.end(), {lr, ur, lc, uc, l64, u64});
        if ((st = launch_split_ldu<double>(c, N, dfr, dfc, dfv, lr, l64, lc, ur, u64, uc))) break;
        if (ic) {
            // exploit the symmetry: U = D L^T from the lower factor (exactly symmetric M)
            std::vector<int> umap, lcol;
            int *dumap = nullptr, *dlcol = nullptr;
            if (!ic_upper_map(N, fr, fc, M->hlrptr, M->hurptr, umap, lcol)) {
                record_error("IC(p): the fill pattern is not symmetric");
                st = SENK_ERR_ARG;
            } else if (!(st = upload(ctx, &dumap, umap.data(), umap.size())) &&
                       !(st = upload(ctx, &dlcol, lcol.data(), lcol.size()))) {
                st = launch_ic_upper(c, M->lnnz, dumap, dlcol, dfv, dpos, l64, u64);
                if (!st && hipStreamSynchronize(c->stream) != hipSuccess) st = SENK_ERR_HIP;
            }
            (void)hipFree(dumap);
            (void)hipFree(dlcol);
            if (st) break;
        }
        v.lrptr = lr; v.lcind = lc; v.l64 = l64;
        v.urptr = ur; v.ucind = uc; v.u64 = u64;
        if (opts->precision & (1 << SENK_FACTOR_F32)) {
            float *l32 = nullptr, *u32 = nullptr;
            if ((st = dalloc(&l32, (size_t)M->lnnz))) break;
            if ((st = dalloc(&u32, (size_t)M->unnz))) break;
            M->owned.insert(M->owned.end(), {l32, u32});
            if ((st = launch_convert<double, float>(c, M->lnnz, l64, l32, nullptr))) break;
            if ((st = launch_convert<double, float>(c, M->unnz, u64, u32, nullptr))) break;
            v.l32 = l32;
            v.u32 = u32;
        }
        // solve layouts; a pattern the layout cannot hold keeps the CSR kernels.  A wave holds
        // 64 / bsize whole blocks: blocks that do not divide 64 (e.g. ABMC blocks of 48) are
        // solved in sub-blocks of gcd(bsize, 64) rows, each level split in order
        v.lts = v.ls;
        v.uts = v.us;
        if (64 % bsize != 0) {
            const int g = std::gcd(bsize, 64);
            std::vector<int> lsub, usub;
            v.lts = split_schedule(v.ls, use_colors ? noblk : lblk, g, lsub);
            v.uts = split_schedule(v.us, use_colors ? noblk : ublk, g, usub);
            if ((st = upload(ctx, &v.lts.dblk, lsub.data(), lsub.size()))) break;
            M->owned.push_back(v.lts.dblk);
            if ((st = upload(ctx, &v.uts.dblk, usub.data(), usub.size()))) break;
            M->owned.push_back(v.uts.dblk);
        }
        if (!opts->csr_only && tri_layout_build(c, v.lts, false, N, lr, lc, &M->lt) == SENK_OK &&
            tri_layout_build(c, v.uts, true, N, ur, uc, &M->ut) == SENK_OK) {
            if ((st = tri_layout_fill<double>(c, v.lts, &M->lt, lr, lc, l64))) break;
            if ((st = tri_layout_fill<double>(c, v.uts, &M->ut, ur, uc, u64))) break;
            if (v.l32) {
                if ((st = tri_layout_fill<float>(c, v.lts, &M->lt, lr, lc, v.l32))) break;
                if ((st = tri_layout_fill<float>(c, v.uts, &M->ut, ur, uc, v.u32))) break;
            }
            if (opts->precision & (1 << SENK_FACTOR_F16)) {
                if ((st = tri_layout_fill_f16(c, v.lts, &M->lt, lr, lc, l64))) break;
                if ((st = tri_layout_fill_f16(c, v.uts, &M->ut, ur, uc, u64))) break;
            }
            v.lt = &M->lt;
            v.ut = &M->ut;
            // ILU(0) with no elimination update of an off-diagonal entry: the split form of
            // the mixed forms' inner BiCGStab
            if (v.l32 && A->v64 && fnnz == nnzA && !(opts->precision & (1 << SENK_FACTOR_F16)) &&
                dilu_pattern(N, fr, fc)) {
                if ((st = dalloc(&M->dilu_d, (size_t)N))) break;
                if ((st = dalloc(&M->dilu_e, (size_t)N))) break;
                M->owned.insert(M->owned.end(), {M->dilu_d, M->dilu_e});
                if ((st = launch_dilu_diag(c, N, ur, u64, A->rptr, A->cind, A->v64, M->dilu_d, M->dilu_e))) break;
                v.dilu_d = M->dilu_d;
                v.dilu_e = M->dilu_e;
                // the split inner CG needs a symmetric preconditioner: IC, or ILU(0) of a
                // symmetric A (U = D L^T to rounding)
                v.symmetric = ic || symmetric_csr(N, rptr, cind, val, kSymTol);
            }
        } else {
            tri_layout_free(&M->lt);
            tri_layout_free(&M->ut);
            // FP16 factors live in the solve layout only
            if (opts->precision & (1 << SENK_FACTOR_F16)) {
                st = SENK_ERR_UNSUPPORTED;
                break;
            }
        }
        if (hipStreamSynchronize(c->stream) != hipSuccess) st = SENK_ERR_HIP;
    } while (0);
    (void)hipFree(dfr);
    (void)hipFree(dfc);
    (void)hipFree(dfv);
    (void)hipFree(dpos);
    std::free(fv);
    std::free(fc);
    std::free(fr);
    if (st) {
        senk_prec_destroy(M);
        return st;
    }
    *out = M;
    return SENK_OK;
}
int senk_prec_destroy(senk_prec M) {
    if (!M) return SENK_ERR_ARG;
    (void)hipStreamSynchronize(M->ctx->c.stream);
    for (void *p : M->owned) (void)hipFree(p);
    senk_amd::tri_layout_free(&M->lt);
    senk_amd::tri_layout_free(&M->ut);
    delete M;
    return SENK_OK;
}
int senk_prec_apply_f64(senk_prec M, const double *dx, double *dy) {
    if (!M) return SENK_ERR_ARG;
    return prec_apply<double>(&M->ctx->c, M->view, dx, nullptr, dy, nullptr, nullptr);
}
int senk_prec_apply_f32(senk_prec M, const float *dx, float *dy) {
    if (!M || !M->view.l32) return SENK_ERR_ARG;
    return prec_apply<float>(&M->ctx->c, M->view, dx, nullptr, dy, nullptr, nullptr);
}
int senk_prec_apply_f16(senk_prec M, const float *dx, float *dy) {
    if (!M || !M->lt.ve16 || !M->ut.ve16) return SENK_ERR_ARG;
    PrecView v = M->view;
    v.f16 = true;
    return prec_apply<float>(&M->ctx->c, v, dx, nullptr, dy, nullptr, nullptr);
}
int senk_prec_set_trsv_mode(senk_prec M, int mode) {
    if (!M || mode < -1 || mode > senk_amd::kTrsvXcd) return SENK_ERR_ARG;
    SENK_HIP(hipStreamSynchronize(M->ctx->c.stream));
    const senk_amd::TriLayout defaults;
    M->lt.mode = M->ut.mode = mode < 0 ? defaults.mode : mode;
    return SENK_OK;
}
int senk_prec_set_split(senk_prec M, int enable) {
    if (!M) return SENK_ERR_ARG;
    if (enable && !M->dilu_d) return SENK_ERR_UNSUPPORTED;
    M->view.dilu_d = enable ? M->dilu_d : nullptr;
    M->view.dilu_e = enable ? M->dilu_e : nullptr;
    return SENK_OK;
}
int senk_prec_check(senk_prec M) {
    if (!M) return SENK_ERR_ARG;
    const int a = senk_amd::tri_layout_check(&M->ctx->c, &M->lt);
    const int b = senk_amd::tri_layout_check(&M->ctx->c, &M->ut);
    return a ? a : b;
}
int senk_prec_factor_nnz(senk_prec M, int *lnnz, int *unnz, int *nlev_l, int *nlev_u) {
    if (!M) return SENK_ERR_ARG;
    if (lnnz) *lnnz = M->lnnz;
    if (unnz) *unnz = M->unnz;
    if (nlev_l) *nlev_l = M->view.ls.nlev;
    if (nlev_u) *nlev_u = M->view.us.nlev;
    return SENK_OK;
}
int senk_prec_download(senk_prec M, int *lrptr, int *lcind, double *lval, int *urptr, int *ucind, double *uval) {
    if (!M) return SENK_ERR_ARG;
    const int N = M->view.N;
    Ctx *c = &M->ctx->c;
    if (lrptr) SENK_HIP(hipMemcpyAsync(lrptr, M->view.lrptr, sizeof(int) * (N + 1), hipMemcpyDeviceToHost, c->stream));
    if (lcind) SENK_HIP(hipMemcpyAsync(lcind, M->view.lcind, sizeof(int) * M->lnnz, hipMemcpyDeviceToHost, c->stream));
    if (lval) SENK_HIP(hipMemcpyAsync(lval, M->view.l64, sizeof(double) * M->lnnz, hipMemcpyDeviceToHost, c->stream));
    if (urptr) SENK_HIP(hipMemcpyAsync(urptr, M->view.urptr, sizeof(int) * (N + 1), hipMemcpyDeviceToHost, c->stream));
    if (ucind) SENK_HIP(hipMemcpyAsync(ucind, M->view.ucind, sizeof(int) * M->unnz, hipMemcpyDeviceToHost, c->stream));
    if (uval) SENK_HIP(hipMemcpyAsync(uval, M->view.u64, sizeof(double) * M->unnz, hipMemcpyDeviceToHost, c->stream));
    SENK_HIP(hipStreamSynchronize(c->stream));
    return SENK_OK;
}

void senk_solve_opts_default(senk_solve_opts *o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->solver = SENK_GMRES;
    o->m = 30;
    o->max_iter = 6000;
    o->eps = 1e-10;
    o->ortho = SENK_MGS;
    o->mixed = 0;
    o->factor_prec = SENK_FACTOR_F64;
    o->inner_eps = 1e-5;
    o->inner_max = 600;
    o->check_every = 0;
    o->outer_m = 20;
}
int senk_solve(senk_ctx ctx, senk_mat A, senk_prec M, const double *db, double *dx, const senk_solve_opts *o,
               senk_result *res, double *hist, int hist_cap) {
    CTX_CHECK(ctx);
    if (!A || !db || !dx || !o || !res) return SENK_ERR_ARG;
    double nrm_b = 0;
    SENK_TRY(senk_nrm2_f64(ctx, db, A->N, &nrm_b));
    PrecView pv;
    pv.N = A->N;
    if (M) pv = M->view;
    LearnedPrec *lp = nullptr;
    if (o->use_learned) {
        if (!M || !M->gnn) return SENK_ERR_ARG;
        // a Krylov method that is not flexible takes the correction only through a mixed form
        if (!o->mixed && o->solver != SENK_FGCR && o->solver != SENK_FGMRES) {
            record_error("use_learned: FGCR / FGMRES, or a mixed form");
            return SENK_ERR_ARG;
        }
        lp = gnn_learned(M->gnn);
        lp->off = false;
        pv.learned = lp;
    }
    if (o->factor_prec == SENK_FACTOR_F16) {
        // FP16 factors serve the FP32 applications of the mixed forms
        if (!o->mixed || !M || !M->lt.ve16 || !M->ut.ve16) return SENK_ERR_ARG;
        pv.f16 = true;
    }
    if (o->mixed < 0 || o->mixed > 3) return SENK_ERR_ARG;
    if (o->mixed == 3) {
        // FP64 Krylov solve whose ILU / ILUB application runs on the FP32 factor copies.
        // Not right-preconditioned GMRES: it forms x from M^-1 (V y) at the end of a
        // cycle, a new application whose FP32 rounding the cycle's residual never saw
        // (true residual ~1e-7 where the recurrence reports 1e-10); the solvers that keep
        // the preconditioned vectors they use (FGMRES, GCR, FGCR, BiCGStab, CG) stay
        // consistent.
        if (!M || o->use_learned || pv.f16 || o->solver == SENK_GMRES) {
            record_error("mixed = 3: needs FP32 factor copies, no learned correction, no FP16, not GMRES");
            return SENK_ERR_ARG;
        }
        if (!(pv.kind == 1 && pv.l32 && pv.u32) && !(pv.kind == 2 && pv.bl32 && pv.bu32)) return SENK_ERR_ARG;
        const size_t ld = ((size_t)A->N + 3) & ~(size_t)3;  // 16-byte aligned halves
        void *buf = nullptr;
        SENK_TRY(ctx_outer(&ctx->c, 2 * sizeof(float) * ld, &buf));
        pv.apply32 = true;
        pv.in32 = (float *)buf;
        pv.out32 = pv.in32 + ld;
    }
    if (o->mixed == 1 || o->mixed == 2) {
        if (!A->v32 || (M && !M->view.l32 && !pv.f16)) return SENK_ERR_ARG;
        if (o->mixed == 2)
            SENK_TRY(nested_solve(&ctx->c, A->view, pv, db, dx, *o, nrm_b, res, hist, hist_cap));
        else
            SENK_TRY(ir_solve(&ctx->c, A->view, pv, db, dx, *o, nrm_b, res, hist, hist_cap));
    } else {
        SolveParams p;
        p.solver = o->solver;
        p.m = o->m;
        p.outer = (o->max_iter + o->m - 1) / o->m;
        p.max_iter = o->max_iter;
        p.eps = o->eps;
        p.nrm_b = nrm_b;
        p.ortho = o->ortho;
        p.check_every = o->check_every;
        p.verbose = o->verbose;
        SENK_TRY(krylov_solve<double>(&ctx->c, A->view, pv, db, dx, p, res, hist, hist_cap));
    }
    double tr = 0;
    SENK_TRY(true_relres(&ctx->c, A->view, db, dx, nrm_b, &tr));
    res->true_relres = tr;
    if (lp) res->fell_back = lp->off ? 1 : 0;
    if (M) SENK_TRY(senk_prec_check(M));
    return res->status;
}

// ---------------------------------------------------------------- learned correction
int senk_gnn_create(senk_ctx ctx, senk_gnn *out) {
    CTX_CHECK(ctx);
    if (!out) return SENK_ERR_ARG;
    return gnn_create(&ctx->c, out);
}
int senk_gnn_destroy(senk_gnn g) {
    if (!g) return SENK_ERR_ARG;
    gnn_destroy(g);
    return SENK_OK;
}
int senk_gnn_set_params(senk_gnn g, const float *theta) {
    if (!g || !theta) return SENK_ERR_ARG;
    return gnn_set_params(g, theta);
}
int senk_gnn_get_params(senk_gnn g, float *theta) {
    if (!g || !theta) return SENK_ERR_ARG;
    return gnn_get_params(g, theta);
}
int senk_gnn_bind(senk_gnn g, senk_mat A) {
    if (!g || !A || !A->v32) return SENK_ERR_ARG;
    return gnn_bind(g, A->N, A->rptr, A->cind, A->v32);
}
int senk_gnn_apply_f32(senk_gnn g, const float *dr, float *dz) {
    if (!g || !dr || !dz) return SENK_ERR_ARG;
    return gnn_apply_f32(g, dr, dz);
}
static PrecView plain_view(senk_prec M) {
    PrecView v = M->view;
    v.learned = nullptr;
    return v;
}
int senk_gnn_loss_grad(senk_gnn g, senk_mat A, senk_prec M, const float *dR, int nvec, double *loss, double *grad) {
    return senk_gnn_loss_grad_pairs(g, A, M, dR, nullptr, nvec, loss, grad);
}
int senk_gnn_loss_grad_pairs(senk_gnn g, senk_mat A, senk_prec M, const float *dR, const float *dZ, int nvec,
                             double *loss, double *grad) {
    if (!g || !A || !M || !dR || nvec <= 0 || (!dZ && !M->view.l32) || !A->v32) return SENK_ERR_ARG;
    return gnn_loss_grad(g, A->view, plain_view(M), dR, dZ, nvec, loss, grad);
}
int senk_gnn_train(senk_gnn g, int nmat, const senk_mat *A, const senk_prec *M, const float *const *dR,
                   const int *nvec, int epochs, int batch, double lr, unsigned seed, double *loss_hist, int verbose) {
    return senk_gnn_train_pairs(g, nmat, A, M, dR, nullptr, nvec, epochs, batch, lr, seed, loss_hist, verbose);
}
int senk_gnn_train_pairs(senk_gnn g, int nmat, const senk_mat *A, const senk_prec *M, const float *const *dR,
                         const float *const *dZ, const int *nvec, int epochs, int batch, double lr, unsigned seed,
                         double *loss_hist, int verbose) {
    if (!g || nmat <= 0 || !A || !M || !dR || !nvec) return SENK_ERR_ARG;
    std::vector<MatView> av((size_t)nmat);
    std::vector<PrecView> pv((size_t)nmat);
    for (int k = 0; k < nmat; k++) {
        if (!A[k] || !M[k] || !dR[k] || nvec[k] <= 0 || A[k]->N != M[k]->view.N || (dZ && !dZ[k]))
            return SENK_ERR_ARG;
        av[k] = A[k]->view;
        pv[k] = plain_view(M[k]);
    }
    return gnn_train(g, nmat, av.data(), pv.data(), dR, dZ, nvec, epochs, batch, lr, seed, loss_hist, verbose);
}
int senk_gnn_collect_nested(senk_ctx ctx, senk_mat A, senk_prec M, const double *db, const senk_solve_opts *opts,
                            float *dR, float *dZ, int cap, int *count) {
    CTX_CHECK(ctx);
    if (!A || !M || !db || !opts || !dR || !dZ || cap <= 0 || !count || opts->mixed != 2 || !A->v32 ||
        opts->solver == SENK_FGCR || opts->solver == SENK_FGMRES)
        return SENK_ERR_ARG;
    double nrm_b = 0;
    SENK_TRY(senk_nrm2_f64(ctx, db, A->N, &nrm_b));
    double *dx = nullptr;
    SENK_TRY(dalloc(&dx, (size_t)A->N));
    SENK_HIP(hipMemsetAsync(dx, 0, sizeof(double) * A->N, ctx->c.stream));
    NestedRecord rec;
    rec.R = dR;
    rec.Z = dZ;
    rec.cap = cap;
    senk_result res;
    const int st = nested_solve(&ctx->c, A->view, plain_view(M), db, dx, *opts, nrm_b, &res, nullptr, 0, &rec);
    (void)hipStreamSynchronize(ctx->c.stream);
    (void)hipFree(dx);
    *count = rec.count;
    return st;
}
int senk_gnn_collect(senk_ctx ctx, senk_mat A, senk_prec M, const double *db, const double *dxstar, int m,
                     int ncycles, double *dR, double *dE, int *count) {
    CTX_CHECK(ctx);
    if (!A || !M || !db || !dR || m <= 0 || ncycles <= 0 || !count) return SENK_ERR_ARG;
    return collect_pairs(&ctx->c, A->view, plain_view(M), db, dxstar, m, ncycles, dR, dE, count);
}
int senk_prec_attach_gnn(senk_prec M, senk_gnn g) {
    if (!M) return SENK_ERR_ARG;
    M->gnn = g;
    return SENK_OK;
}

}  // extern "C"
