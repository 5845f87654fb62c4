// senk_main: the reference's end-to-end workflow (main.cpp:9-187) on the device, with
// the same steps, parameters and printed lines:
//
//   read a MatrixMarket file (zeros removed) -> CheckStructure -> Expand ("L", symmetric
//   files) or Duplicate -> Padding to a multiple of 128 -> ABMC ordering (blocks of 128,
//   "connect") and Reordering -> symmetric Jacobi scaling (D^-1/2 A D^-1/2) -> the ILU
//   pattern closed under 8x1 blocks (AllocBlockZero) -> Ilu0 (on the device, over the
//   pattern's dependency levels: the same factor as the sequential one) -> Split "L-DU"
//   with the inverted diagonal -> b ~ U[-1,1] from std::mt19937(1), zero on the padding
//   rows, x0 = 0 -> IluGmresm with m = 50, 6000/50 restarts, epsilon 1e-8, timed in
//   milliseconds as the reference times it -> "# test <relative residual>".
//
// Usage: senk_main <matrix.mtx>   (the reference prefixes a fixed directory to argv[1];
// here the argument is the path itself).  The unused BCSR copies of the factors that
// main.cpp:111-112 builds are not built.  Errors print the library's message and exit 1.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "senk_amd.h"
#include "senk_gpu.hpp"

namespace {

void need(int st, const char *what) {
    if (st != SENK_OK) {
        std::printf("%s: %s (%s)\n", what, senk_status_string(st), senk_last_error());
        std::exit(1);
    }
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: %s <matrix.mtx>\n", argv[0]);
        return 2;
    }
    const char *filename = argv[1];

    // io::ReadMatrixMarket (senk_io.hpp:45,68 print these two lines)
    std::printf("# Readfile MatrixMarket\n");
    int N, M, shape;
    long long entries;
    need(senk_matrix_market_size(filename, &N, &M, &entries, &shape), "ReadMatrixMarket");
    std::printf("# %d %d %lld\n", N, M, entries);
    double *tval;
    int *tcind, *trptr;
    need(senk_read_matrix_market(filename, &tval, &tcind, &trptr, &N, &M, &shape, 1), "ReadMatrixMarket");
    std::printf("# %d\n", trptr[N]);
    if (!senk_check_structure(tcind, trptr, N)) {
        std::printf("False: CheckStructure\n");
        std::exit(1);
    }
    double *val;
    int *cind, *rptr;
    if (shape == SENK_SHAPE_SYM) need(senk_expand(tval, tcind, trptr, &val, &cind, &rptr, N, "L"), "Expand");
    else need(senk_duplicate(tval, tcind, trptr, &val, &cind, &rptr, N), "Duplicate");
    senk_host_free(tval);
    senk_host_free(tcind);
    senk_host_free(trptr);

    // Padding
    const int ori_N = N;
    need(senk_padding(&val, &cind, &rptr, 128, &N), "Padding");
    std::printf("%d\n", rptr[N]);
    M = N;
    std::printf("%d %d\n", N, M);

    // Reordering
    int num_color;
    int *size_color, *LP, *RP;
    need(senk_abmc_permutation(cind, rptr, &num_color, &size_color, &LP, &RP, N, 128, 0, "connect"),
         "GetABMCPermutation");
    need(senk_reordering(val, cind, rptr, LP, RP, N), "Reordering");
    std::printf("# ");
    for (int i = 0; i < num_color; i++) std::printf("%d ", size_color[i + 1] - size_color[i]);
    std::printf("\n");
    std::printf("# Reordered!\n");

    // Scaling
    double *tdiag;
    need(senk_get_diag(val, cind, rptr, &tdiag, N), "GetDiag");
    for (int i = 0; i < N; i++) tdiag[i] = 1 / std::sqrt(std::abs(tdiag[i]));
    need(senk_scaling(val, cind, rptr, tdiag, tdiag, N), "Scaling");

    // ILU factorization
    need(senk_duplicate(val, cind, rptr, &tval, &tcind, &trptr, N), "Duplicate");
    const int bnl = 8, bnw = 1;
    need(senk_alloc_block_zero(&tval, &tcind, &trptr, N, bnl, bnw), "AllocBlockZero");
    std::printf("# Allocated!\n");
    {
        using senk::gpu_detail::Dev;
        const size_t fnnz = (size_t)trptr[N];
        Dev<double> dv(tval, fnnz);
        Dev<int> dc(tcind, fnnz), dr(trptr, (size_t)N + 1);
        need(senk_ilu0_f64(senk::gpu_detail::ctx(), dv.p, dc.p, dr.p, N, nullptr, 0, 0, trptr, tcind), "Ilu0");
        dv.to_host(tval);
    }
    std::printf("# Factrized!\n");

    double *lval, *uval;
    int *lcind, *lrptr, *ucind, *urptr;
    const int diagInv = 1;
    need(senk_split(tval, tcind, trptr, &lval, &lcind, &lrptr, &uval, &ucind, &urptr, nullptr, N, "L-DU", diagInv),
         "Split");
    std::printf("# Splited!\n");

    std::vector<double> x(N, 0.0), b(N);
    std::mt19937 engine(1);
    std::uniform_real_distribution<> dist1(-1.0, 1.0);
    for (int i = 0; i < N; i++) b[i] = dist1(engine);
    for (int i = 0; i < N - ori_N; i++) b[ori_N + i] = 0.0;

    const int max_iter = 6000;
    const double epsilon = 1.0e-8;
    double nrm_b = 0;
    for (int i = 0; i < N; i++) nrm_b += b[i] * b[i];
    nrm_b = std::sqrt(nrm_b);

    auto start = std::chrono::system_clock::now();
    senk::solver::IluGmresm<double>(val, cind, rptr, lval, lcind, lrptr, uval, ucind, urptr, b.data(), x.data(), nrm_b,
                                    max_iter / 50, 50, N, epsilon);
    auto end = std::chrono::system_clock::now();
    const double elapsed = (double)std::chrono::duration_cast<std::chrono::milliseconds>(end - start).count();
    std::printf("%e\n", elapsed);

    // test::RelativeResidualError (senk_test.hpp:11-22), in FP64 on the host
    double nt = 0, nb = 0;
    for (int i = 0; i < N; i++) {
        double t = 0;
        for (int j = rptr[i]; j < rptr[i + 1]; j++) t += val[j] * x[cind[j]];
        t = b[i] - t;
        nt += t * t;
        nb += b[i] * b[i];
    }
    std::printf("# test %e\n", std::sqrt(nt) / std::sqrt(nb));

    for (void *p : {(void *)val, (void *)cind, (void *)rptr, (void *)size_color, (void *)LP, (void *)RP,
                    (void *)tdiag, (void *)tval, (void *)tcind, (void *)trptr, (void *)lval, (void *)lcind,
                    (void *)lrptr, (void *)uval, (void *)ucind, (void *)urptr})
        senk_host_free(p);
    return 0;
}
