// The relative residual ||b - A x|| / ||b|| (test::RelativeResidualError,
// senk_test.hpp:11-22) with no allocation at all: its scratch -- the per-workgroup partial
// sums of ||b - A x||^2 and ||b||^2 -- lives in a file-scope device array instead of the
// context's reduction slots or a work vector, and the result stays on the device.  For
// callers that cannot allocate (code under stream capture, tight serving loops).
//
// Every workgroup writes its own partials before the finishing kernel reads them, so the
// array needs no clearing; two calls that overlap on one device (two streams, two host
// threads) race on it.
#include "dev_common.hpp"
#include "solvers.hpp"

#include <mutex>

namespace senk_amd {
constexpr int kNaBlocks = kRedMaxBlocks;  // workgroups of the partial-sum pass (1024)
constexpr int kNaLanes = 8;               // lanes per row, as the solvers' residual pass
constexpr int kMaxDevices = 64;

// [0, kNaBlocks): ||r||^2 partials; [kNaBlocks, 2 kNaBlocks): ||b||^2 partials.
// 16 KiB per device, held until the process ends.  In host code this name is not a device
// address: kernels receive the address hipGetSymbolAddress returns on each device.
static __device__ __align__(256) double scratch[2 * kNaBlocks];

namespace {
std::mutex g_mu;
double *g_scratch[kMaxDevices] = {};  // the array's address on each device

// r_i = 1*b_i + (-1)*(A x)_i with the row's sum split over kNaLanes lanes (the solvers'
// residual pass, k_residual), squares summed per workgroup; b_i^2 likewise.
__global__ __launch_bounds__(kBlock) void k_na_relres_part(int N, const int *__restrict__ rptr,
                                                           const int *__restrict__ cind,
                                                           const double *__restrict__ val,
                                                           const double *__restrict__ x,
                                                           const double *__restrict__ b, double *part) {
    const int rows_per_pass = (int)((long long)gridDim.x * kBlock / kNaLanes);
    const long long gid = (long long)blockIdx.x * kBlock + threadIdx.x;
    const int q = (int)(gid % kNaLanes);
    double rr = 0, bb = 0;
    for (int row = (int)(gid / kNaLanes); row < N; row += rows_per_pass) {
        double acc = 0;
        const int e = rptr[row + 1];
        for (int j = rptr[row] + q; j < e; j += kNaLanes) acc += val[j] * x[cind[j]];
#pragma unroll
        for (int off = kNaLanes / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, kNaLanes);
        if (q == 0) {
            const double bi = b[row];
            const double r = 1.0 * bi + (-1.0) * acc;
            rr += r * r;
            bb += bi * bi;
        }
    }
    const double v[2] = {rr, bb};
    store_partials<2>(v, part);
}
__global__ __launch_bounds__(kBlock) void k_na_relres_fin(const double *part, int B, double *out) {
    const double rr = sum_partials(part, B), bb = sum_partials(part + kPartLd, B);
    if (threadIdx.x == 0) *out = sqrt(rr) / sqrt(bb);
}
}  // namespace

// Looks up the array's address on `device` (current on entry): called once per device by
// senk_ctx_create.
int noalloc_init(int device) {
    if (device < 0 || device >= kMaxDevices) return SENK_ERR_ARG;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[device]) return SENK_OK;
    void *p = nullptr;
    SENK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch[device] = static_cast<double *>(p);
    return SENK_OK;
}

int relres_noalloc(Ctx *c, const double *dval, const int *dcind, const int *drptr, const double *db,
                   const double *dx, int N, double *drelres) {
    double *part = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (c->device >= 0 && c->device < kMaxDevices) part = g_scratch[c->device];
    }
    if (!part) {
        record_error("no-allocation residual: the context's device has no scratch address");
        return SENK_ERR_ARG;
    }
    long long threads = (long long)N * kNaLanes;
    int grid = (int)((threads + kBlock - 1) / kBlock);
    if (grid > kNaBlocks) grid = kNaBlocks;
    k_na_relres_part<<<grid, kBlock, 0, c->stream>>>(N, drptr, dcind, dval, dx, db, part);
    SENK_LAUNCH_CHECK();
    k_na_relres_fin<<<1, kBlock, 0, c->stream>>>(part, grid, drelres);
    SENK_LAUNCH_CHECK();
    return SENK_OK;
}

}  // namespace senk_amd
