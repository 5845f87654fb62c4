// Test infrastructure only: a C-ABI shim over the reference SenK headers, compiled
// from the headers where they lie (include path given by oracle/Makefile) into
// oracle/_ref/libsenkref.so.  Nothing in the product links or loads this file; the
// parity tests and bench.py's cpu_baseline leg load it to pin the C oracle
// (oracle/senk_oracle.c) and to time the reference's own OpenMP path.
//
// Every function forwards to one reference template.  Functions whose reference
// counterpart allocates outputs (T** arguments) copy them into caller buffers of a
// stated capacity and free the reference's allocation; they return the produced
// length, or -(needed length) when the capacity is too small.
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cmath>

#include "senk.hpp"
#include "senk_test.hpp"

#define EXPORT extern "C" __attribute__((visibility("default")))

namespace {
template <typename T>
int copy_out(T *src, T *dst, int n, int cap) {
    if (n > cap) { free(src); return -n; }
    if (n > 0) std::memcpy(dst, src, sizeof(T) * (size_t)n);
    free(src);
    return n;
}
}  // namespace

// ---------------------------------------------------------------- SpMV / SpTRSV
EXPORT void ref_spmv_csr_d(double *val, int *cind, int *rptr, double *x, double *y, int N) {
    senk::sparse::SpmvCsr<double>(val, cind, rptr, x, y, N);
}
EXPORT void ref_spmv_csr_f(float *val, int *cind, int *rptr, float *x, float *y, int N) {
    senk::sparse::SpmvCsr<float>(val, cind, rptr, x, y, N);
}
EXPORT void ref_spmv_csr_diag_d(double *val, int *cind, int *rptr, double *diag, double *x, double *y, int N) {
    senk::sparse::SpmvCsr<double>(val, cind, rptr, diag, x, y, N);
}
EXPORT void ref_spmv_sell_d(double *val, int *cind, int *wid, int len, double *x, double *y, int N) {
    senk::sparse::SpmvSell<double>(val, cind, wid, len, x, y, N);
}
EXPORT void ref_spmv_sell_f(float *val, int *cind, int *wid, int len, float *x, float *y, int N) {
    senk::sparse::SpmvSell<float>(val, cind, wid, len, x, y, N);
}

#define BCSR_DISPATCH(BNL, BNW, CALL)                                   \
    if (BNL == 2 && BNW == 1) { constexpr int bl = 2, bw = 1; CALL; return 0; } \
    if (BNL == 2 && BNW == 2) { constexpr int bl = 2, bw = 2; CALL; return 0; } \
    if (BNL == 3 && BNW == 3) { constexpr int bl = 3, bw = 3; CALL; return 0; } \
    if (BNL == 4 && BNW == 1) { constexpr int bl = 4, bw = 1; CALL; return 0; } \
    if (BNL == 4 && BNW == 2) { constexpr int bl = 4, bw = 2; CALL; return 0; } \
    if (BNL == 4 && BNW == 4) { constexpr int bl = 4, bw = 4; CALL; return 0; } \
    if (BNL == 8 && BNW == 1) { constexpr int bl = 8, bw = 1; CALL; return 0; } \
    return -1;

EXPORT int ref_spmv_bcsr_d(double *bval, int *bcind, int *brptr, double *x, double *y, int N, int bnl, int bnw) {
    BCSR_DISPATCH(bnl, bnw, (senk::sparse::SpmvBcsr<double, bl, bw>(bval, bcind, brptr, x, y, N)))
}
EXPORT int ref_spmv_bcsr_f(float *bval, int *bcind, int *brptr, float *x, float *y, int N, int bnl, int bnw) {
    BCSR_DISPATCH(bnl, bnw, (senk::sparse::SpmvBcsr<float, bl, bw>(bval, bcind, brptr, x, y, N)))
}
EXPORT void ref_sptrsv_l_d(double *val, int *cind, int *rptr, double *x, double *y, int N) {
    senk::sparse::SptrsvCsr_l<double>(val, cind, rptr, x, y, N);
}
EXPORT void ref_sptrsv_u_d(double *val, int *cind, int *rptr, double *x, double *y, int N) {
    senk::sparse::SptrsvCsr_u<double>(val, cind, rptr, x, y, N);
}
EXPORT void ref_sptrsv_l_f(float *val, int *cind, int *rptr, float *x, float *y, int N) {
    senk::sparse::SptrsvCsr_l<float>(val, cind, rptr, x, y, N);
}
EXPORT void ref_sptrsv_u_f(float *val, int *cind, int *rptr, float *x, float *y, int N) {
    senk::sparse::SptrsvCsr_u<float>(val, cind, rptr, x, y, N);
}
EXPORT void ref_sptrsv_l_amc_d(double *val, int *cind, int *rptr, double *x, double *y, int N, int *cptr, int cnum) {
    senk::sparse::SptrsvCsr_l<double>(val, cind, rptr, x, y, N, cptr, cnum);
}
EXPORT void ref_sptrsv_u_amc_d(double *val, int *cind, int *rptr, double *x, double *y, int N, int *cptr, int cnum) {
    senk::sparse::SptrsvCsr_u<double>(val, cind, rptr, x, y, N, cptr, cnum);
}
EXPORT void ref_sptrsv_l_abmc_d(double *val, int *cind, int *rptr, double *x, double *y, int N, int *cptr, int cnum, int bsize) {
    senk::sparse::SptrsvCsr_l<double>(val, cind, rptr, x, y, N, cptr, cnum, bsize);
}
EXPORT void ref_sptrsv_u_abmc_d(double *val, int *cind, int *rptr, double *x, double *y, int N, int *cptr, int cnum, int bsize) {
    senk::sparse::SptrsvCsr_u<double>(val, cind, rptr, x, y, N, cptr, cnum, bsize);
}
EXPORT void ref_sptrsv_l_bj_d(double *val, int *cind, int *rptr, double *x, double *y, int N, int bnum) {
    senk::sparse::SptrsvCsr_l<double>(val, cind, rptr, x, y, N, bnum);
}
EXPORT void ref_sptrsv_u_bj_d(double *val, int *cind, int *rptr, double *x, double *y, int N, int bnum) {
    senk::sparse::SptrsvCsr_u<double>(val, cind, rptr, x, y, N, bnum);
}
EXPORT int ref_sptrsv_bcsr_l_d(double *bval, int *bcind, int *brptr, double *x, double *y, int N, int bnl, int bnw) {
    BCSR_DISPATCH(bnl, bnw, (senk::sparse::SptrsvBcsr_l<double, bl, bw>(bval, bcind, brptr, x, y, N)))
}
EXPORT int ref_sptrsv_bcsr_u_d(double *bval, int *bcind, int *brptr, double *x, double *y, int N, int bnl, int bnw) {
    BCSR_DISPATCH(bnl, bnw, (senk::sparse::SptrsvBcsr_u<double, bl, bw>(bval, bcind, brptr, x, y, N)))
}
EXPORT int ref_sptrsv_bcsr_l_abmc_d(double *bval, int *bcind, int *brptr, double *x, double *y, int N,
                                    int *cptr, int cnum, int bsize, int bnl, int bnw) {
    BCSR_DISPATCH(bnl, bnw, (senk::sparse::SptrsvBcsr_l<double, bl, bw>(bval, bcind, brptr, x, y, N, cptr, cnum, bsize)))
}
EXPORT int ref_sptrsv_bcsr_u_abmc_d(double *bval, int *bcind, int *brptr, double *x, double *y, int N,
                                    int *cptr, int cnum, int bsize, int bnl, int bnw) {
    BCSR_DISPATCH(bnl, bnw, (senk::sparse::SptrsvBcsr_u<double, bl, bw>(bval, bcind, brptr, x, y, N, cptr, cnum, bsize)))
}

// ---------------------------------------------------------------- BLAS-1 / BLAS-2
EXPORT void ref_copy_d(double *x, double *y, int N) { senk::blas1::Copy<double>(x, y, N); }
EXPORT void ref_scal_d(double a, double *x, int N) { senk::blas1::Scal<double>(a, x, N); }
EXPORT void ref_axpy_d(double a, double *x, double *y, int N) { senk::blas1::Axpy<double>(a, x, y, N); }
EXPORT void ref_axpby_d(double a, double *x, double b, double *y, int N) { senk::blas1::Axpby<double>(a, x, b, y, N); }
EXPORT void ref_axpyz_d(double a, double *x, double *y, double *z, int N) { senk::blas1::Axpyz<double>(a, x, y, z, N); }
EXPORT double ref_dot_d(double *x, double *y, int N) { return senk::blas1::Dot<double>(x, y, N); }
EXPORT double ref_nrm2_d(double *x, int N) { return senk::blas1::Nrm2<double>(x, N); }
EXPORT float ref_dot_f(float *x, float *y, int N) { return senk::blas1::Dot<float>(x, y, N); }
EXPORT float ref_nrm2_f(float *x, int N) { return senk::blas1::Nrm2<float>(x, N); }
EXPORT void ref_hadprod_d(double *x, double *y, int N) { senk::blas1::HadProd<double>(x, y, N); }
EXPORT void ref_haddiv_d(double *x, double *y, int N) { senk::blas1::HadDiv<double>(x, y, N); }
EXPORT double ref_ggen_d(double a, double b, double *c, double *s) { return senk::blas1::Ggen<double>(a, b, c, s); }
EXPORT void ref_grot_d(double c, double s, double *a, double *b) { senk::blas1::Grot<double>(c, s, a, b); }
EXPORT void ref_trsv_d(double *U, double *b, double *x, int n, int m) { senk::blas2::Trsv<double>(U, b, x, n, m); }
EXPORT void ref_convert_d2f(double *in, float *out, int n) { senk::utils::Convert<double, float>(in, out, n); }

// ---------------------------------------------------------------- matrix setup
EXPORT int ref_check_structure_d(double *val, int *cind, int *rptr, int N) {
    return senk::matrix::CheckStructure<double>(val, cind, rptr, N) ? 1 : 0;
}
EXPORT void ref_csr2csc_d(double *val, int *cind, int *rptr, double *cval, int *crind, int *ccptr, int N, int M) {
    double *v; int *r, *c;
    senk::matrix::Csr2Csc<double>(val, cind, rptr, &v, &r, &c, N, M);
    int nnz = rptr[N];
    std::memcpy(cval, v, sizeof(double) * nnz);
    std::memcpy(crind, r, sizeof(int) * nnz);
    std::memcpy(ccptr, c, sizeof(int) * (M + 1));
    free(v); free(r); free(c);
}
// s_len must hold num_slice+1 ints; s_val/s_cind capacity cap.
EXPORT int ref_csr2sell_d(double *val, int *cind, int *rptr, double *s_val, int *s_cind, int *s_len,
                          int size, int N, int cap) {
    double *v; int *c; int *l;
    int num_slice = (N + size - 1) / size;
    int nnz = senk::matrix::Csr2Sell<double>(val, cind, rptr, &v, &c, &l, size, N);
    std::memcpy(s_len, l, sizeof(int) * (num_slice + 1));
    free(l);
    int r = copy_out(v, s_val, nnz, cap);
    copy_out(c, s_cind, nnz, cap);
    return r;
}
// Returns the number of blocks; bval needs cap_blocks*bnl*bnw.
EXPORT int ref_csr2bcsr_d(double *val, int *cind, int *rptr, double *bval, int *bcind, int *brptr,
                          int N, int bnl, int bnw, int cap_blocks) {
    double *v; int *c, *r;
    senk::matrix::Csr2Bcsr<double>(val, cind, rptr, &v, &c, &r, N, bnl, bnw);
    int nb = r[N / bnl];
    std::memcpy(brptr, r, sizeof(int) * (N / bnl + 1));
    free(r);
    if (nb > cap_blocks) { free(v); free(c); return -nb; }
    std::memcpy(bval, v, sizeof(double) * (size_t)nb * bnl * bnw);
    std::memcpy(bcind, c, sizeof(int) * nb);
    free(v); free(c);
    return nb;
}
EXPORT int ref_bcsr2csr_d(double *bval, int *bcind, int *brptr, double *val, int *cind, int *rptr,
                          int N, int bnl, int bnw, int cap) {
    double *v; int *c, *r;
    int nnz = senk::matrix::Bcsr2Csr<double>(bval, bcind, brptr, &v, &c, &r, N, bnl, bnw);
    std::memcpy(rptr, r, sizeof(int) * (N + 1));
    free(r);
    int res = copy_out(v, val, nnz, cap);
    copy_out(c, cind, nnz, cap);
    return res;
}
// In/out arrays must already have room for the padded size (N rounded up to size).
// io::ReadMatrixMarket: dims = {N, M, shape (0 Sym, 1 Unsym)}; returns nnz, -(nnz + 1)
// when a capacity is short, INT_MIN when the reader refused the file.
EXPORT int ref_read_mm(const char *fn, int removeZeros, double *val, int *cind, int *rptr, int *dims, int cap,
                       int rcap) {
    double *v; int *c, *r; int N = 0, M = 0;
    senk::io::Shape sh;
    if (!senk::io::ReadMatrixMarket(fn, &v, &c, &r, &N, &M, &sh, removeZeros != 0)) return INT_MIN;
    dims[0] = N; dims[1] = M; dims[2] = sh == senk::io::Sym ? 0 : 1;
    const int nnz = r[N];
    if (nnz > cap || N + 1 > rcap) { free(v); free(c); free(r); return -(nnz + 1); }
    std::memcpy(val, v, sizeof(double) * nnz);
    std::memcpy(cind, c, sizeof(int) * nnz);
    std::memcpy(rptr, r, sizeof(int) * (N + 1));
    free(v); free(c); free(r);
    return nnz;
}
// Expand (key "L" or "U"): returns the entry count, -(count + 1) when cap is short.
EXPORT int ref_expand_d(double *tval, int *tcind, int *trptr, int N, const char *key, double *val, int *cind,
                        int *rptr, int cap) {
    double *v; int *c, *r;
    senk::matrix::Expand<double>(tval, tcind, trptr, &v, &c, &r, N, key);
    const int nnz = r[N];
    if (nnz > cap) { free(v); free(c); free(r); return -(nnz + 1); }
    std::memcpy(val, v, sizeof(double) * nnz);
    std::memcpy(cind, c, sizeof(int) * nnz);
    std::memcpy(rptr, r, sizeof(int) * (N + 1));
    free(v); free(c); free(r);
    return nnz;
}
// RemoveZeros on malloc'd copies; outputs into the caller's arrays (same capacity as
// the input); returns the new entry count.
EXPORT int ref_remove_zeros_d(double *val, int *cind, int *rptr, int N) {
    const int nnz = rptr[N];
    double *v = (double *)malloc(sizeof(double) * (nnz ? nnz : 1)); std::memcpy(v, val, sizeof(double) * nnz);
    int *c = (int *)malloc(sizeof(int) * (nnz ? nnz : 1)); std::memcpy(c, cind, sizeof(int) * nnz);
    senk::matrix::RemoveZeros<double>(&v, &c, rptr, N);
    std::memcpy(val, v, sizeof(double) * rptr[N]);
    std::memcpy(cind, c, sizeof(int) * rptr[N]);
    free(v); free(c);
    return rptr[N];
}
EXPORT int ref_padding_d(double *val, int *cind, int *rptr, int size, int N) {
    int nnz = rptr[N];
    double *v = (double *)malloc(sizeof(double) * nnz); std::memcpy(v, val, sizeof(double) * nnz);
    int *c = (int *)malloc(sizeof(int) * nnz); std::memcpy(c, cind, sizeof(int) * nnz);
    int *r = (int *)malloc(sizeof(int) * (N + 1)); std::memcpy(r, rptr, sizeof(int) * (N + 1));
    int n = N;
    senk::matrix::Padding<double>(&v, &c, &r, size, &n);
    std::memcpy(val, v, sizeof(double) * r[n]);
    std::memcpy(cind, c, sizeof(int) * r[n]);
    std::memcpy(rptr, r, sizeof(int) * (n + 1));
    free(v); free(c); free(r);
    return n;
}
// key: 0 "LD-U", 1 "L-DU", 2 "L-D-U", 3 "LU-D".  Output arrays sized nnz (+N+1 ptrs, N diag).
EXPORT void ref_split_d(double *val, int *cind, int *rptr,
                        double *lval, int *lcind, int *lrptr, double *uval, int *ucind, int *urptr,
                        double *diag, int N, int key, int invDiag) {
    static const char *keys[] = {"LD-U", "L-DU", "L-D-U", "LU-D"};
    double *lv = nullptr, *uv = nullptr, *dv = nullptr; int *lc = nullptr, *uc = nullptr, *lr = nullptr, *ur = nullptr;
    senk::matrix::Split<double>(val, cind, rptr, &lv, &lc, &lr, &uv, &uc, &ur, &dv, N, keys[key], invDiag != 0);
    int lnnz = lr[N];
    std::memcpy(lrptr, lr, sizeof(int) * (N + 1));
    if (lnnz) { std::memcpy(lval, lv, sizeof(double) * lnnz); std::memcpy(lcind, lc, sizeof(int) * lnnz); }
    if (key != 3) {
        int unnz = ur[N];
        std::memcpy(urptr, ur, sizeof(int) * (N + 1));
        if (unnz) { std::memcpy(uval, uv, sizeof(double) * unnz); std::memcpy(ucind, uc, sizeof(int) * unnz); }
    }
    if ((key == 2 || key == 3) && diag && dv) std::memcpy(diag, dv, sizeof(double) * N);
    free(lv); free(lc); free(lr);
    if (key != 3) { free(uv); free(uc); free(ur); }
    if (key == 2 || key == 3) free(dv);
}
EXPORT void ref_get_diag_d(double *val, int *cind, int *rptr, double *diag, int N) {
    double *d;
    senk::matrix::GetDiag<double>(val, cind, rptr, &d, N);
    std::memcpy(diag, d, sizeof(double) * N);
    free(d);
}
EXPORT void ref_scaling_d(double *val, int *cind, int *rptr, double *ld, double *rd, int N) {
    senk::matrix::Scaling<double>(val, cind, rptr, ld, rd, N);
}
EXPORT void ref_ilu0_d(double *val, int *cind, int *rptr, int N) { senk::matrix::Ilu0<double>(val, cind, rptr, N); }
EXPORT void ref_ilu0_f(float *val, int *cind, int *rptr, int N) { senk::matrix::Ilu0<float>(val, cind, rptr, N); }

static int pattern_out(double *v, int *c, int *r, double *val, int *cind, int *rptr, int N, int cap) {
    int nnz = r[N];
    std::memcpy(rptr, r, sizeof(int) * (N + 1));
    free(r);
    int res = copy_out(v, val, nnz, cap);
    copy_out(c, cind, nnz, cap);
    return res;
}
static void dup_in(double *val, int *cind, int *rptr, int N, double **v, int **c, int **r) {
    int nnz = rptr[N];
    *v = (double *)malloc(sizeof(double) * nnz); std::memcpy(*v, val, sizeof(double) * nnz);
    *c = (int *)malloc(sizeof(int) * nnz); std::memcpy(*c, cind, sizeof(int) * nnz);
    *r = (int *)malloc(sizeof(int) * (N + 1)); std::memcpy(*r, rptr, sizeof(int) * (N + 1));
}
EXPORT int ref_ilup_d(double *val, int *cind, int *rptr, double *oval, int *ocind, int *orptr, int N, int p, int cap) {
    double *v; int *c, *r;
    dup_in(val, cind, rptr, N, &v, &c, &r);
    senk::matrix::Ilup<double>(&v, &c, &r, N, p);
    return pattern_out(v, c, r, oval, ocind, orptr, N, cap);
}
EXPORT int ref_alloc_level_zero_d(double *val, int *cind, int *rptr, double *oval, int *ocind, int *orptr,
                                  int N, int p, int cap) {
    double *v; int *c, *r;
    dup_in(val, cind, rptr, N, &v, &c, &r);
    senk::matrix::AllocLevelZero<double>(&v, &c, &r, N, p);
    return pattern_out(v, c, r, oval, ocind, orptr, N, cap);
}
EXPORT int ref_alloc_block_zero_d(double *val, int *cind, int *rptr, double *oval, int *ocind, int *orptr,
                                  int N, int bnl, int bnw, int cap) {
    double *v; int *c, *r;
    dup_in(val, cind, rptr, N, &v, &c, &r);
    senk::matrix::AllocBlockZero<double>(&v, &c, &r, N, bnl, bnw);
    return pattern_out(v, c, r, oval, ocind, orptr, N, cap);
}
EXPORT int ref_remove_off_diagonal_d(double *val, int *cind, int *rptr, int N, int bnum) {
    double *v; int *c, *r;
    dup_in(val, cind, rptr, N, &v, &c, &r);
    senk::matrix::RemoveOffDiagonal<double>(&v, &c, &r, N, bnum);
    int nnz = r[N];
    std::memcpy(val, v, sizeof(double) * nnz);
    std::memcpy(cind, c, sizeof(int) * nnz);
    std::memcpy(rptr, r, sizeof(int) * (N + 1));
    free(v); free(c); free(r);
    return nnz;
}
EXPORT void ref_reordering_rp_d(double *val, int *cind, int *rptr, int *RP, int N) {
    senk::matrix::Reordering<double>(val, cind, rptr, RP, N);
}
EXPORT void ref_reordering_d(double *val, int *cind, int *rptr, int *LP, int *RP, int N) {
    senk::matrix::Reordering<double>(val, cind, rptr, LP, RP, N);
}

// ---------------------------------------------------------------- graph ordering
// size_color must hold N+1 ints (the reference allocates N+1 too).
EXPORT int ref_amc_perm(int *cind, int *rptr, int *size_color, int *LP, int *RP, int N, int isSym) {
    int nc; int *sc, *lp, *rp;
    senk::graph::GetAMCPermutation(cind, rptr, &nc, &sc, &lp, &rp, N, isSym != 0);
    std::memcpy(size_color, sc, sizeof(int) * (nc + 1));
    std::memcpy(LP, lp, sizeof(int) * N);
    std::memcpy(RP, rp, sizeof(int) * N);
    free(sc); free(lp); free(rp);
    return nc;
}
// method: 0 "simple", 1 "connect"
EXPORT int ref_abmc_perm(int *cind, int *rptr, int *size_color, int *LP, int *RP, int N, int bsize, int isSym, int method) {
    int nc; int *sc, *lp, *rp;
    senk::graph::GetABMCPermutation(cind, rptr, &nc, &sc, &lp, &rp, N, bsize, isSym != 0,
                                    method ? "connect" : "simple");
    std::memcpy(size_color, sc, sizeof(int) * (nc + 1));
    std::memcpy(LP, lp, sizeof(int) * N);
    std::memcpy(RP, rp, sizeof(int) * N);
    free(sc); free(lp); free(rp);
    return nc;
}

// ---------------------------------------------------------------- solvers
EXPORT void ref_gmresm_d(double *val, int *cind, int *rptr, double *b, double *x, double nrm_b,
                         int outer, int m, int N, double eps) {
    senk::solver::Gmresm<double>(val, cind, rptr, b, x, nrm_b, outer, m, N, eps);
}
EXPORT void ref_ilu_gmresm_d(double *val, int *cind, int *rptr, double *lval, int *lcind, int *lrptr,
                             double *uval, int *ucind, int *urptr, double *b, double *x, double nrm_b,
                             int outer, int m, int N, double eps) {
    senk::solver::IluGmresm<double>(val, cind, rptr, lval, lcind, lrptr, uval, ucind, urptr, b, x, nrm_b, outer, m, N, eps);
}
EXPORT void ref_ilu_gmresm_f(float *val, int *cind, int *rptr, float *lval, int *lcind, int *lrptr,
                             float *uval, int *ucind, int *urptr, float *b, float *x, float nrm_b,
                             int outer, int m, int N, float eps) {
    senk::solver::IluGmresm<float>(val, cind, rptr, lval, lcind, lrptr, uval, ucind, urptr, b, x, nrm_b, outer, m, N, eps);
}
EXPORT void ref_amc_ilu_gmresm_d(double *val, int *cind, int *rptr, double *lval, int *lcind, int *lrptr,
                                 double *uval, int *ucind, int *urptr, int *cptr, int cnum,
                                 double *b, double *x, double nrm_b, int outer, int m, int N, double eps) {
    senk::solver::AmcIluGmresm<double>(val, cind, rptr, lval, lcind, lrptr, uval, ucind, urptr, cptr, cnum,
                                       b, x, nrm_b, outer, m, N, eps);
}
EXPORT void ref_abmc_ilu_gmresm_d(double *val, int *cind, int *rptr, double *lval, int *lcind, int *lrptr,
                                  double *uval, int *ucind, int *urptr, int *cptr, int cnum, int bsize,
                                  double *b, double *x, double nrm_b, int outer, int m, int N, double eps) {
    senk::solver::AbmcIluGmresm<double>(val, cind, rptr, lval, lcind, lrptr, uval, ucind, urptr, cptr, cnum, bsize,
                                        b, x, nrm_b, outer, m, N, eps);
}
EXPORT void ref_bj_ilu_gmresm_d(double *val, int *cind, int *rptr, double *lval, int *lcind, int *lrptr,
                                double *uval, int *ucind, int *urptr, int bnum,
                                double *b, double *x, double nrm_b, int outer, int m, int N, double eps) {
    senk::solver::BjIluGmresm<double>(val, cind, rptr, lval, lcind, lrptr, uval, ucind, urptr, bnum,
                                      b, x, nrm_b, outer, m, N, eps);
}
EXPORT int ref_ilub_gmresm_d(double *val, int *cind, int *rptr, double *blval, int *blcind, int *blrptr,
                             double *buval, int *bucind, int *burptr, double *b, double *x, double nrm_b,
                             int outer, int m, int N, double eps, int bnl, int bnw) {
    BCSR_DISPATCH(bnl, bnw, (senk::solver::IlubGmresm<double, bl, bw>(val, cind, rptr, blval, blcind, blrptr,
                                                                   buval, bucind, burptr, b, x, nrm_b, outer, m, N, eps)))
}
EXPORT int ref_abmc_ilub_gmresm_d(double *val, int *cind, int *rptr, double *blval, int *blcind, int *blrptr,
                                  double *buval, int *bucind, int *burptr, int *cptr, int cnum, int bsize,
                                  double *b, double *x, double nrm_b, int outer, int m, int N, double eps,
                                  int bnl, int bnw) {
    BCSR_DISPATCH(bnl, bnw, (senk::solver::AbmcIlubGmresm<double, bl, bw>(val, cind, rptr, blval, blcind, blrptr,
                                                                       buval, bucind, burptr, cptr, cnum, bsize,
                                                                       b, x, nrm_b, outer, m, N, eps)))
}
EXPORT int ref_bcsr_gmresm_d(double *bval, int *bcind, int *brptr, double *b, double *x, double nrm_b,
                             int outer, int m, int N, double eps, int bnl, int bnw) {
    BCSR_DISPATCH(bnl, bnw, (senk::solver::Gmresm<double, bl, bw>(bval, bcind, brptr, b, x, nrm_b, outer, m, N, eps)))
}
EXPORT void ref_gcrm_d(double *val, int *cind, int *rptr, double *b, double *x, double nrm_b,
                       int outer, int m, int N, double eps) {
    senk::solver::Gcrm<double>(val, cind, rptr, b, x, nrm_b, outer, m, N, eps);
}
EXPORT void ref_ilu_gcrm_d(double *val, int *cind, int *rptr, double *lval, int *lcind, int *lrptr,
                           double *uval, int *ucind, int *urptr, double *b, double *x, double nrm_b,
                           int outer, int m, int N, double eps) {
    senk::solver::IluGcrm<double>(val, cind, rptr, lval, lcind, lrptr, uval, ucind, urptr, b, x, nrm_b, outer, m, N, eps);
}
EXPORT int ref_ilub_gcrm_d(double *val, int *cind, int *rptr, double *blval, int *blcind, int *blrptr,
                           double *buval, int *bucind, int *burptr, double *b, double *x, double nrm_b,
                           int outer, int m, int N, double eps, int bnl, int bnw) {
    BCSR_DISPATCH(bnl, bnw, (senk::solver::IlubGcrm<double, bl, bw>(val, cind, rptr, blval, blcind, blrptr,
                                                                 buval, bucind, burptr, b, x, nrm_b, outer, m, N, eps)))
}
EXPORT void ref_bicgstab_d(double *val, int *cind, int *rptr, double *b, double *x, double nrm_b,
                           int max_iter, int N, double eps) {
    senk::solver::Bicgstab<double>(val, cind, rptr, b, x, nrm_b, max_iter, N, eps);
}
EXPORT void ref_ilu_bicgstab_d(double *val, int *cind, int *rptr, double *lval, int *lcind, int *lrptr,
                               double *uval, int *ucind, int *urptr, double *b, double *x, double nrm_b,
                               int max_iter, int N, double eps) {
    senk::solver::IluBicgstab<double>(val, cind, rptr, lval, lcind, lrptr, uval, ucind, urptr, b, x, nrm_b,
                                      max_iter, N, eps);
}
EXPORT int ref_ilub_bicgstab_d(double *val, int *cind, int *rptr, double *blval, int *blcind, int *blrptr,
                               double *buval, int *bucind, int *burptr, double *b, double *x, double nrm_b,
                               int max_iter, int N, double eps, int bnl, int bnw) {
    BCSR_DISPATCH(bnl, bnw, (senk::solver::IlubBicgstab<double, bl, bw>(val, cind, rptr, blval, blcind, blrptr,
                                                                     buval, bucind, burptr, b, x, nrm_b,
                                                                     max_iter, N, eps)))
}
EXPORT void ref_relative_residual_error(double *val, int *cind, int *rptr, double *b, double *x, int N) {
    senk::test::RelativeResidualError(val, cind, rptr, b, x, N);
    fflush(stdout);
}
EXPORT void ref_flush(void) { fflush(stdout); }
