// C++ drop-in for SenK's preconditioned Krylov path: the solver drivers (senk_gmres.hpp,
// senk_gcr.hpp, senk_bicgstab.hpp), the sparse kernels (senk_sparse.hpp), BLAS-1 and
// BLAS-2 (senk_blas1.hpp, senk_blas2.hpp) and test::RelativeResidualError
// (senk_test.hpp), with the same namespaces, template names, argument order and meaning,
// running on the GPU through the C ABI of senk_amd.h.  Link with -lsenk_amd.
//
// Use: include this header BEFORE any of the reference's headers.  It takes over the
// include guards of the seven headers it replaces, so a later `#include "senk.hpp"` (or
// of the setup headers senk_io / senk_matrix / senk_utils / senk_graph) brings in only
// the host setup code, and the solver and kernel calls resolve to the templates below.
// Including it after one of the replaced headers is a compile error.
//
// Host-pointer mode: every call uploads its arrays, runs on the device and copies the
// outputs back (x of a solver, y of a kernel), so a single call pays the PCIe transfers;
// the handle API of senk_amd.h keeps data resident for repeated solves.  The solvers
// print the lines the reference prints.  Errors print the library's message and call
// exit(1), as the reference does.  One device context per thread, created on first use.
//
// Precision.  T = double runs in double everywhere.  T = float runs the kernels that
// have an FP32 device path in float (SpMV CSR / BCSR / SELL, every SpTRSV form,
// Copy / Scal / Axpy / Axpby / Axpyz / HadProd / HadDiv); Dot and Nrm2 accumulate the
// float products in double and return the sum rounded to float; the solver drivers,
// SpmvCsr with a separate diagonal, Ggen / Grot and Trsv compute in double and round
// their outputs to float once.
#pragma once

#if defined(SENK_BLAS1_HPP) || defined(SENK_BLAS2_HPP) || defined(SENK_SPARSE_HPP) || defined(SENK_GMRES_HPP) || \
    defined(SENK_GCR_HPP) || defined(SENK_BICGSTAB_HPP) || defined(SENK_TEST_HPP)
#error "senk_gpu.hpp replaces senk_blas1/blas2/sparse/gmres/gcr/bicgstab/test.hpp: include it before senk.hpp"
#endif
#define SENK_BLAS1_HPP
#define SENK_BLAS2_HPP
#define SENK_SPARSE_HPP
#define SENK_GMRES_HPP
#define SENK_GCR_HPP
#define SENK_BICGSTAB_HPP
#define SENK_TEST_HPP

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <type_traits>
#include <vector>

#include "senk_amd.h"

namespace senk {
namespace gpu_detail {

inline void check(int st, const char *what) {
    if (st != SENK_OK) {
        std::printf("%s: %s (%s)\n", what, senk_status_string(st), senk_last_error());
        std::exit(1);
    }
}
inline senk_ctx ctx() {
    static thread_local senk_ctx c = nullptr;
    if (!c) check(senk_ctx_create(0, &c), "senk_ctx_create");
    return c;
}
// A device copy of a host array of count elements of type D (converted from H).
template <typename D>
struct Dev {
    D *p = nullptr;
    size_t n = 0;
    Dev(size_t count) : n(count) { check(senk_device_malloc(ctx(), sizeof(D) * (n ? n : 1), (void **)&p), "senk_device_malloc"); }
    template <typename H>
    Dev(const H *h, size_t count) : Dev(count) {
        if (!n) return;
        if (std::is_same<H, D>::value) {
            check(senk_memcpy_htod(ctx(), p, h, sizeof(D) * n), "senk_memcpy_htod");
        } else {
            std::vector<D> tmp(h, h + n);
            check(senk_memcpy_htod(ctx(), p, tmp.data(), sizeof(D) * n), "senk_memcpy_htod");
        }
    }
    template <typename H>
    void to_host(H *h) const {
        if (!n) return;
        if (std::is_same<H, D>::value) {
            check(senk_memcpy_dtoh(ctx(), h, p, sizeof(D) * n), "senk_memcpy_dtoh");
        } else {
            std::vector<D> tmp(n);
            check(senk_memcpy_dtoh(ctx(), tmp.data(), p, sizeof(D) * n), "senk_memcpy_dtoh");
            for (size_t i = 0; i < n; i++) h[i] = (H)tmp[i];
        }
    }
    ~Dev() { (void)senk_device_free(ctx(), p); }
    Dev(const Dev &) = delete;
    Dev &operator=(const Dev &) = delete;
};
// The device value type a kernel runs in: T itself where an FP32 path exists.
template <typename T>
using Val = typename std::conditional<std::is_same<T, float>::value, float, double>::type;
template <typename D>
struct Csr {
    Dev<D> val;
    Dev<int> cind, rptr;
    template <typename T>
    Csr(const T *v, const int *c, const int *r, int N) : val(v, (size_t)r[N]), cind(c, (size_t)r[N]), rptr(r, (size_t)N + 1) {}
};
template <typename D>
struct Bcsr {
    Dev<D> val;
    Dev<int> cind, rptr;
    template <typename T>
    Bcsr(const T *v, const int *c, const int *r, int N, int bnl, int bnw)
        : val(v, (size_t)r[N / bnl] * bnl * bnw), cind(c, (size_t)r[N / bnl]), rptr(r, (size_t)N / bnl + 1) {}
};
// Dependency-level plan of a factor pattern (host arrays), released at scope exit.
struct Plan {
    senk_trsv_plan p = nullptr;
    Plan(const int *rptr, const int *cind, int N, int upper) {
        check(senk_trsv_plan_create(ctx(), rptr, cind, N, upper, &p), "senk_trsv_plan_create");
    }
    ~Plan() { senk_trsv_plan_destroy(p); }
    Plan(const Plan &) = delete;
    Plan &operator=(const Plan &) = delete;
};

// FP64 / FP32 dispatch of the kernel entry points.
inline int spmv_csr(const double *v, const int *c, const int *r, const double *x, double *y, int N) {
    return senk_spmv_csr_f64(ctx(), v, c, r, x, y, N);
}
inline int spmv_csr(const float *v, const int *c, const int *r, const float *x, float *y, int N) {
    return senk_spmv_csr_f32(ctx(), v, c, r, x, y, N);
}
inline int spmv_bcsr(const double *v, const int *c, const int *r, const double *x, double *y, int N, int bnl, int bnw) {
    return senk_spmv_bcsr_f64(ctx(), v, c, r, x, y, N, bnl, bnw);
}
inline int spmv_bcsr(const float *v, const int *c, const int *r, const float *x, float *y, int N, int bnl, int bnw) {
    return senk_spmv_bcsr_f32(ctx(), v, c, r, x, y, N, bnl, bnw);
}
inline int spmv_sell(const double *v, const int *c, const int *w, int len, const double *x, double *y, int N) {
    return senk_spmv_sell_f64(ctx(), v, c, w, len, x, y, N);
}
inline int spmv_sell(const float *v, const int *c, const int *w, int len, const float *x, float *y, int N) {
    return senk_spmv_sell_f32(ctx(), v, c, w, len, x, y, N);
}
inline int trsv_seq(senk_trsv_plan p, bool up, const double *v, const int *c, const int *r, const double *x, double *y,
                    int N) {
    return up ? senk_sptrsv_csr_u_f64(ctx(), p, v, c, r, x, y, N) : senk_sptrsv_csr_l_f64(ctx(), p, v, c, r, x, y, N);
}
inline int trsv_seq(senk_trsv_plan p, bool up, const float *v, const int *c, const int *r, const float *x, float *y,
                    int N) {
    return up ? senk_sptrsv_csr_u_f32(ctx(), p, v, c, r, x, y, N) : senk_sptrsv_csr_l_f32(ctx(), p, v, c, r, x, y, N);
}
inline int trsv_amc(bool up, const double *v, const int *c, const int *r, const double *x, double *y, int N,
                    const int *cptr, int cnum) {
    return up ? senk_sptrsv_csr_u_amc_f64(ctx(), v, c, r, x, y, N, cptr, cnum)
              : senk_sptrsv_csr_l_amc_f64(ctx(), v, c, r, x, y, N, cptr, cnum);
}
inline int trsv_amc(bool up, const float *v, const int *c, const int *r, const float *x, float *y, int N,
                    const int *cptr, int cnum) {
    return up ? senk_sptrsv_csr_u_amc_f32(ctx(), v, c, r, x, y, N, cptr, cnum)
              : senk_sptrsv_csr_l_amc_f32(ctx(), v, c, r, x, y, N, cptr, cnum);
}
inline int trsv_abmc(bool up, const double *v, const int *c, const int *r, const double *x, double *y, int N,
                     const int *cptr, int cnum, int bsize) {
    return up ? senk_sptrsv_csr_u_abmc_f64(ctx(), v, c, r, x, y, N, cptr, cnum, bsize)
              : senk_sptrsv_csr_l_abmc_f64(ctx(), v, c, r, x, y, N, cptr, cnum, bsize);
}
inline int trsv_abmc(bool up, const float *v, const int *c, const int *r, const float *x, float *y, int N,
                     const int *cptr, int cnum, int bsize) {
    return up ? senk_sptrsv_csr_u_abmc_f32(ctx(), v, c, r, x, y, N, cptr, cnum, bsize)
              : senk_sptrsv_csr_l_abmc_f32(ctx(), v, c, r, x, y, N, cptr, cnum, bsize);
}
inline int trsv_bj(senk_trsv_plan p, bool up, const double *v, const int *c, const int *r, const double *x, double *y,
                   int N, int bnum) {
    return up ? senk_sptrsv_csr_u_bj_f64(ctx(), p, v, c, r, x, y, N, bnum)
              : senk_sptrsv_csr_l_bj_f64(ctx(), p, v, c, r, x, y, N, bnum);
}
inline int trsv_bj(senk_trsv_plan p, bool up, const float *v, const int *c, const int *r, const float *x, float *y,
                   int N, int bnum) {
    return up ? senk_sptrsv_csr_u_bj_f32(ctx(), p, v, c, r, x, y, N, bnum)
              : senk_sptrsv_csr_l_bj_f32(ctx(), p, v, c, r, x, y, N, bnum);
}
inline int trsv_bcsr(bool up, const double *v, const int *c, const int *r, const double *x, double *y, int N, int bnl,
                     int bnw) {
    return up ? senk_sptrsv_bcsr_u_f64(ctx(), v, c, r, x, y, N, bnl, bnw)
              : senk_sptrsv_bcsr_l_f64(ctx(), v, c, r, x, y, N, bnl, bnw);
}
inline int trsv_bcsr(bool up, const float *v, const int *c, const int *r, const float *x, float *y, int N, int bnl,
                     int bnw) {
    return up ? senk_sptrsv_bcsr_u_f32(ctx(), v, c, r, x, y, N, bnl, bnw)
              : senk_sptrsv_bcsr_l_f32(ctx(), v, c, r, x, y, N, bnl, bnw);
}
inline int trsv_bcsr_abmc(bool up, const double *v, const int *c, const int *r, const double *x, double *y, int N,
                          const int *cptr, int cnum, int bsize, int bnl, int bnw) {
    return up ? senk_sptrsv_bcsr_u_abmc_f64(ctx(), v, c, r, x, y, N, cptr, cnum, bsize, bnl, bnw)
              : senk_sptrsv_bcsr_l_abmc_f64(ctx(), v, c, r, x, y, N, cptr, cnum, bsize, bnl, bnw);
}
inline int trsv_bcsr_abmc(bool up, const float *v, const int *c, const int *r, const float *x, float *y, int N,
                          const int *cptr, int cnum, int bsize, int bnl, int bnw) {
    return up ? senk_sptrsv_bcsr_u_abmc_f32(ctx(), v, c, r, x, y, N, cptr, cnum, bsize, bnl, bnw)
              : senk_sptrsv_bcsr_l_abmc_f32(ctx(), v, c, r, x, y, N, cptr, cnum, bsize, bnl, bnw);
}
inline int copy(const double *x, double *y, int N) { return senk_copy_f64(ctx(), x, y, N); }
inline int copy(const float *x, float *y, int N) { return senk_copy_f32(ctx(), x, y, N); }
inline int scal(double a, double *x, int N) { return senk_scal_f64(ctx(), a, x, N); }
inline int scal(float a, float *x, int N) { return senk_scal_f32(ctx(), a, x, N); }
inline int axpy(double a, const double *x, double *y, int N) { return senk_axpy_f64(ctx(), a, x, y, N); }
inline int axpy(float a, const float *x, float *y, int N) { return senk_axpy_f32(ctx(), a, x, y, N); }
inline int axpby(double a, const double *x, double b, double *y, int N) { return senk_axpby_f64(ctx(), a, x, b, y, N); }
inline int axpby(float a, const float *x, float b, float *y, int N) { return senk_axpby_f32(ctx(), a, x, b, y, N); }
inline int axpyz(double a, const double *x, const double *y, double *z, int N) {
    return senk_axpyz_f64(ctx(), a, x, y, z, N);
}
inline int axpyz(float a, const float *x, const float *y, float *z, int N) {
    return senk_axpyz_f32(ctx(), a, x, y, z, N);
}
inline int dot(const double *x, const double *y, int N, double *r) { return senk_dot_f64(ctx(), x, y, N, r); }
inline int dot(const float *x, const float *y, int N, double *r) { return senk_dot_f32(ctx(), x, y, N, r); }
inline int hadprod(const double *x, double *y, int N) { return senk_hadprod_f64(ctx(), x, y, N); }
inline int hadprod(const float *x, float *y, int N) { return senk_hadprod_f32(ctx(), x, y, N); }
inline int haddiv(const double *x, double *y, int N) { return senk_haddiv_f64(ctx(), x, y, N); }
inline int haddiv(const float *x, float *y, int N) { return senk_haddiv_f32(ctx(), x, y, N); }
}  // namespace gpu_detail

namespace solver {

// senk_gmres.hpp:33-95
template <typename T>
void Gmresm(T *val, int *cind, int *rptr, T *b, T *x, T nrm_b, int outer, int m, int N, T epsilon) {
    using namespace gpu_detail;
    Csr<double> A(val, cind, rptr, N);
    Dev<double> db(b, N), dx(x, N);
    check(senk_gmresm_f64(ctx(), A.val.p, A.cind.p, A.rptr.p, db.p, dx.p, nrm_b, outer, m, N, epsilon, 1, nullptr,
                          nullptr, 0), "Gmresm");
    dx.to_host(x);
}
// senk_gmres.hpp:113-175 (BCSR matrix)
template <typename T, int bnl, int bnw>
void Gmresm(T *bval, int *bcind, int *brptr, T *b, T *x, T nrm_b, int outer, int m, int N, T epsilon) {
    using namespace gpu_detail;
    Bcsr<double> A(bval, bcind, brptr, N, bnl, bnw);
    Dev<double> db(b, N), dx(x, N);
    check(senk_gmresm_bcsr_f64(ctx(), A.val.p, A.cind.p, A.rptr.p, db.p, dx.p, nrm_b, outer, m, N, epsilon, bnl,
                               bnw, 1, nullptr, nullptr, 0), "Gmresm");
    dx.to_host(x);
}
// senk_gmres.hpp:197-270 (sequential SpTRSV: run on the factors' dependency levels)
template <typename T>
void IluGmresm(T *val, int *cind, int *rptr, T *lval, int *lcind, int *lrptr, T *uval, int *ucind, int *urptr,
               T *b, T *x, T nrm_b, int outer, int m, int N, T epsilon) {
    using namespace gpu_detail;
    Csr<double> A(val, cind, rptr, N), L(lval, lcind, lrptr, N), U(uval, ucind, urptr, N);
    Dev<double> db(b, N), dx(x, N);
    Plan lp(lrptr, lcind, N, 0), up(urptr, ucind, N, 1);
    check(senk_ilu_gmresm_f64(ctx(), A.val.p, A.cind.p, A.rptr.p, L.val.p, L.cind.p, L.rptr.p, U.val.p, U.cind.p,
                              U.rptr.p, lp.p, up.p, db.p, dx.p, nrm_b, outer, m, N, epsilon, 1, nullptr, nullptr, 0),
          "IluGmresm");
    dx.to_host(x);
}
// senk_gmres.hpp:293-367
template <typename T>
void AmcIluGmresm(T *val, int *cind, int *rptr, T *lval, int *lcind, int *lrptr, T *uval, int *ucind, int *urptr,
                  int *cptr, int cnum, T *b, T *x, T nrm_b, int outer, int m, int N, T epsilon) {
    using namespace gpu_detail;
    Csr<double> A(val, cind, rptr, N), L(lval, lcind, lrptr, N), U(uval, ucind, urptr, N);
    Dev<double> db(b, N), dx(x, N);
    check(senk_amc_ilu_gmresm_f64(ctx(), A.val.p, A.cind.p, A.rptr.p, L.val.p, L.cind.p, L.rptr.p, U.val.p,
                                  U.cind.p, U.rptr.p, cptr, cnum, db.p, dx.p, nrm_b, outer, m, N, epsilon, 1, nullptr,
                                  nullptr, 0), "AmcIluGmresm");
    dx.to_host(x);
}
// senk_gmres.hpp:391-465
template <typename T>
void AbmcIluGmresm(T *val, int *cind, int *rptr, T *lval, int *lcind, int *lrptr, T *uval, int *ucind,
                   int *urptr, int *cptr, int cnum, int bsize, T *b, T *x, T nrm_b, int outer, int m, int N,
                   T epsilon) {
    using namespace gpu_detail;
    Csr<double> A(val, cind, rptr, N), L(lval, lcind, lrptr, N), U(uval, ucind, urptr, N);
    Dev<double> db(b, N), dx(x, N);
    check(senk_abmc_ilu_gmresm_f64(ctx(), A.val.p, A.cind.p, A.rptr.p, L.val.p, L.cind.p, L.rptr.p, U.val.p,
                                   U.cind.p, U.rptr.p, cptr, cnum, bsize, db.p, dx.p, nrm_b, outer, m, N, epsilon, 1,
                                   nullptr, nullptr, 0), "AbmcIluGmresm");
    dx.to_host(x);
}
// senk_gmres.hpp:487-561
template <typename T>
void BjIluGmresm(T *val, int *cind, int *rptr, T *lval, int *lcind, int *lrptr, T *uval, int *ucind, int *urptr,
                 int bnum, T *b, T *x, T nrm_b, int outer, int m, int N, T epsilon) {
    using namespace gpu_detail;
    Csr<double> A(val, cind, rptr, N), L(lval, lcind, lrptr, N), U(uval, ucind, urptr, N);
    Dev<double> db(b, N), dx(x, N);
    Plan lp(lrptr, lcind, N, 0), up(urptr, ucind, N, 1);
    check(senk_bj_ilu_gmresm_f64(ctx(), A.val.p, A.cind.p, A.rptr.p, L.val.p, L.cind.p, L.rptr.p, U.val.p, U.cind.p,
                                 U.rptr.p, lp.p, up.p, bnum, db.p, dx.p, nrm_b, outer, m, N, epsilon, 1, nullptr,
                                 nullptr, 0), "BjIluGmresm");
    dx.to_host(x);
}
// senk_gmres.hpp:582-655
template <typename T, int bnl, int bnw>
void IlubGmresm(T *val, int *cind, int *rptr, T *blval, int *blcind, int *blrptr, T *buval, int *bucind,
                int *burptr, T *b, T *x, T nrm_b, int outer, int m, int N, T epsilon) {
    using namespace gpu_detail;
    Csr<double> A(val, cind, rptr, N);
    Bcsr<double> L(blval, blcind, blrptr, N, bnl, bnw), U(buval, bucind, burptr, N, bnl, bnw);
    Dev<double> db(b, N), dx(x, N);
    check(senk_ilub_gmresm_f64(ctx(), A.val.p, A.cind.p, A.rptr.p, L.val.p, L.cind.p, L.rptr.p, U.val.p, U.cind.p,
                               U.rptr.p, db.p, dx.p, nrm_b, outer, m, N, epsilon, bnl, bnw, 1, nullptr, nullptr, 0),
          "IlubGmresm");
    dx.to_host(x);
}
// senk_gmres.hpp:679-753
template <typename T, int bnl, int bnw>
void AbmcIlubGmresm(T *val, int *cind, int *rptr, T *blval, int *blcind, int *blrptr, T *buval, int *bucind,
                    int *burptr, int *cptr, int cnum, int bsize, T *b, T *x, T nrm_b, int outer, int m, int N,
                    T epsilon) {
    using namespace gpu_detail;
    Csr<double> A(val, cind, rptr, N);
    Bcsr<double> L(blval, blcind, blrptr, N, bnl, bnw), U(buval, bucind, burptr, N, bnl, bnw);
    Dev<double> db(b, N), dx(x, N);
    check(senk_abmc_ilub_gmresm_f64(ctx(), A.val.p, A.cind.p, A.rptr.p, L.val.p, L.cind.p, L.rptr.p, U.val.p,
                                    U.cind.p, U.rptr.p, cptr, cnum, bsize, db.p, dx.p, nrm_b, outer, m, N, epsilon,
                                    bnl, bnw, 1, nullptr, nullptr, 0), "AbmcIlubGmresm");
    dx.to_host(x);
}
// senk_gcr.hpp:30-72
template <typename T>
void Gcrm(T *val, int *cind, int *rptr, T *b, T *x, T nrm_b, int outer, int m, int N, T epsilon) {
    using namespace gpu_detail;
    Csr<double> A(val, cind, rptr, N);
    Dev<double> db(b, N), dx(x, N);
    check(senk_gcrm_f64(ctx(), A.val.p, A.cind.p, A.rptr.p, db.p, dx.p, nrm_b, outer, m, N, epsilon, 1, nullptr,
                        nullptr, 0), "Gcrm");
    dx.to_host(x);
}
// senk_gcr.hpp:93-142
template <typename T>
void IluGcrm(T *val, int *cind, int *rptr, T *lval, int *lcind, int *lrptr, T *uval, int *ucind, int *urptr, T *b,
             T *x, T nrm_b, int outer, int m, int N, T epsilon) {
    using namespace gpu_detail;
    Csr<double> A(val, cind, rptr, N), L(lval, lcind, lrptr, N), U(uval, ucind, urptr, N);
    Dev<double> db(b, N), dx(x, N);
    Plan lp(lrptr, lcind, N, 0), up(urptr, ucind, N, 1);
    check(senk_ilu_gcrm_f64(ctx(), A.val.p, A.cind.p, A.rptr.p, L.val.p, L.cind.p, L.rptr.p, U.val.p, U.cind.p,
                            U.rptr.p, lp.p, up.p, db.p, dx.p, nrm_b, outer, m, N, epsilon, 1, nullptr, nullptr, 0),
          "IluGcrm");
    dx.to_host(x);
}
// senk_gcr.hpp:163-212
template <typename T, int bnl, int bnw>
void IlubGcrm(T *val, int *cind, int *rptr, T *blval, int *blcind, int *blrptr, T *buval, int *bucind, int *burptr,
              T *b, T *x, T nrm_b, int outer, int m, int N, T epsilon) {
    using namespace gpu_detail;
    Csr<double> A(val, cind, rptr, N);
    Bcsr<double> L(blval, blcind, blrptr, N, bnl, bnw), U(buval, bucind, burptr, N, bnl, bnw);
    Dev<double> db(b, N), dx(x, N);
    check(senk_ilub_gcrm_f64(ctx(), A.val.p, A.cind.p, A.rptr.p, L.val.p, L.cind.p, L.rptr.p, U.val.p, U.cind.p,
                             U.rptr.p, db.p, dx.p, nrm_b, outer, m, N, epsilon, bnl, bnw, 1, nullptr, nullptr, 0),
          "IlubGcrm");
    dx.to_host(x);
}
// senk_bicgstab.hpp:29-87
template <typename T>
void Bicgstab(T *val, int *cind, int *rptr, T *b, T *x, T nrm_b, int max_iter, int N, T epsilon) {
    using namespace gpu_detail;
    Csr<double> A(val, cind, rptr, N);
    Dev<double> db(b, N), dx(x, N);
    check(senk_bicgstab_f64(ctx(), A.val.p, A.cind.p, A.rptr.p, db.p, dx.p, nrm_b, max_iter, N, epsilon, 1, nullptr,
                            nullptr, 0), "Bicgstab");
    dx.to_host(x);
}
// senk_bicgstab.hpp:107-176
template <typename T>
void IluBicgstab(T *val, int *cind, int *rptr, T *lval, int *lcind, int *lrptr, T *uval, int *ucind, int *urptr,
                 T *b, T *x, T nrm_b, int max_iter, int N, T epsilon) {
    using namespace gpu_detail;
    Csr<double> A(val, cind, rptr, N), L(lval, lcind, lrptr, N), U(uval, ucind, urptr, N);
    Dev<double> db(b, N), dx(x, N);
    Plan lp(lrptr, lcind, N, 0), up(urptr, ucind, N, 1);
    check(senk_ilu_bicgstab_f64(ctx(), A.val.p, A.cind.p, A.rptr.p, L.val.p, L.cind.p, L.rptr.p, U.val.p, U.cind.p,
                                U.rptr.p, lp.p, up.p, db.p, dx.p, nrm_b, max_iter, N, epsilon, 1, nullptr, nullptr, 0),
          "IluBicgstab");
    dx.to_host(x);
}
// senk_bicgstab.hpp:198-267
template <typename T, int bnl, int bnw>
void IlubBicgstab(T *val, int *cind, int *rptr, T *blval, int *blcind, int *blrptr, T *buval, int *bucind,
                  int *burptr, T *b, T *x, T nrm_b, int max_iter, int N, T epsilon) {
    using namespace gpu_detail;
    Csr<double> A(val, cind, rptr, N);
    Bcsr<double> L(blval, blcind, blrptr, N, bnl, bnw), U(buval, bucind, burptr, N, bnl, bnw);
    Dev<double> db(b, N), dx(x, N);
    check(senk_ilub_bicgstab_f64(ctx(), A.val.p, A.cind.p, A.rptr.p, L.val.p, L.cind.p, L.rptr.p, U.val.p, U.cind.p,
                                 U.rptr.p, db.p, dx.p, nrm_b, max_iter, N, epsilon, bnl, bnw, 1, nullptr, nullptr, 0),
          "IlubBicgstab");
    dx.to_host(x);
}

}  // namespace solver

namespace sparse {

// senk_sparse.hpp:25-35
template <typename T>
inline void SpmvCsr(T *val, int *cind, int *rptr, T *x, T *y, int N) {
    using namespace gpu_detail;
    Csr<Val<T>> A(val, cind, rptr, N);
    Dev<Val<T>> dx(x, N), dy(N);
    check(spmv_csr(A.val.p, A.cind.p, A.rptr.p, dx.p, dy.p, N), "SpmvCsr");
    dy.to_host(y);
}
// senk_sparse.hpp:47-57 (CSR without the diagonal plus a separate diagonal)
template <typename T>
inline void SpmvCsr(T *val, int *cind, int *rptr, T *diag, T *x, T *y, int N) {
    using namespace gpu_detail;
    Csr<double> A(val, cind, rptr, N);
    Dev<double> dd(diag, N), dx(x, N), dy(N);
    check(senk_spmv_csr_diag_f64(ctx(), A.val.p, A.cind.p, A.rptr.p, dd.p, dx.p, dy.p, N), "SpmvCsr");
    dy.to_host(y);
}
// senk_sparse.hpp:70-91
template <typename T, int bnl, int bnw>
inline void SpmvBcsr(T *bval, int *bcind, int *brptr, T *x, T *y, int N) {
    using namespace gpu_detail;
    Bcsr<Val<T>> A(bval, bcind, brptr, N, bnl, bnw);
    Dev<Val<T>> dx(x, N), dy(N);
    check(spmv_bcsr(A.val.p, A.cind.p, A.rptr.p, dx.p, dy.p, N, bnl, bnw), "SpmvBcsr");
    dy.to_host(y);
}
// senk_sparse.hpp:103-121, on the layout Csr2Sell writes (senk_matrix.hpp:90-142)
template <typename T>
inline void SpmvSell(T *val, int *cind, int *wid, int len, T *x, T *y, int N) {
    using namespace gpu_detail;
    const int ns = (N + len - 1) / len;
    const size_t nnz = (size_t)wid[ns] * len;
    Dev<Val<T>> dv(val, nnz), dx(x, N), dy(N);
    Dev<int> dc(cind, nnz), dw(wid, (size_t)ns + 1);
    check(spmv_sell(dv.p, dc.p, dw.p, len, dx.p, dy.p, N), "SpmvSell");
    dy.to_host(y);
}
// senk_sparse.hpp:132-143 (sequential; run on the factor's dependency levels)
template <typename T>
inline void SptrsvCsr_l(T *val, int *cind, int *rptr, T *x, T *y, int N) {
    using namespace gpu_detail;
    Csr<Val<T>> L(val, cind, rptr, N);
    Dev<Val<T>> dx(x, N), dy(N);
    Plan p(rptr, cind, N, 0);
    check(trsv_seq(p.p, false, L.val.p, L.cind.p, L.rptr.p, dx.p, dy.p, N), "SptrsvCsr_l");
    dy.to_host(y);
}
// senk_sparse.hpp:154-167 (x may alias y)
template <typename T>
inline void SptrsvCsr_u(T *val, int *cind, int *rptr, T *x, T *y, int N) {
    using namespace gpu_detail;
    Csr<Val<T>> U(val, cind, rptr, N);
    Dev<Val<T>> dx(x, N), dy(N);
    Plan p(rptr, cind, N, 1);
    check(trsv_seq(p.p, true, U.val.p, U.cind.p, U.rptr.p, dx.p, dy.p, N), "SptrsvCsr_u");
    dy.to_host(y);
}
// senk_sparse.hpp:180-201 (AMC: cptr counts rows)
template <typename T>
inline void SptrsvCsr_l(T *val, int *cind, int *rptr, T *x, T *y, int N, int *cptr, int cnum) {
    using namespace gpu_detail;
    Csr<Val<T>> L(val, cind, rptr, N);
    Dev<Val<T>> dx(x, N), dy(N);
    check(trsv_amc(false, L.val.p, L.cind.p, L.rptr.p, dx.p, dy.p, N, cptr, cnum), "SptrsvCsr_l");
    dy.to_host(y);
}
// senk_sparse.hpp:214-236
template <typename T>
inline void SptrsvCsr_u(T *val, int *cind, int *rptr, T *x, T *y, int N, int *cptr, int cnum) {
    using namespace gpu_detail;
    Csr<Val<T>> U(val, cind, rptr, N);
    Dev<Val<T>> dx(x, N), dy(N);
    check(trsv_amc(true, U.val.p, U.cind.p, U.rptr.p, dx.p, dy.p, N, cptr, cnum), "SptrsvCsr_u");
    dy.to_host(y);
}
// senk_sparse.hpp:250-275 (ABMC: cptr counts blocks of bsize rows)
template <typename T>
inline void SptrsvCsr_l(T *val, int *cind, int *rptr, T *x, T *y, int N, int *cptr, int cnum, int bsize) {
    using namespace gpu_detail;
    Csr<Val<T>> L(val, cind, rptr, N);
    Dev<Val<T>> dx(x, N), dy(N);
    check(trsv_abmc(false, L.val.p, L.cind.p, L.rptr.p, dx.p, dy.p, N, cptr, cnum, bsize), "SptrsvCsr_l");
    dy.to_host(y);
}
// senk_sparse.hpp:289-315
template <typename T>
inline void SptrsvCsr_u(T *val, int *cind, int *rptr, T *x, T *y, int N, int *cptr, int cnum, int bsize) {
    using namespace gpu_detail;
    Csr<Val<T>> U(val, cind, rptr, N);
    Dev<Val<T>> dx(x, N), dy(N);
    check(trsv_abmc(true, U.val.p, U.cind.p, U.rptr.p, dx.p, dy.p, N, cptr, cnum, bsize), "SptrsvCsr_u");
    dy.to_host(y);
}
// senk_sparse.hpp:328-346 (block-Jacobi: bnum ranges of N/bnum rows, RemoveOffDiagonal factors)
template <typename T>
inline void SptrsvCsr_l(T *val, int *cind, int *rptr, T *x, T *y, int N, int bnum) {
    using namespace gpu_detail;
    Csr<Val<T>> L(val, cind, rptr, N);
    Dev<Val<T>> dx(x, N), dy(N);
    Plan p(rptr, cind, N, 0);
    check(trsv_bj(p.p, false, L.val.p, L.cind.p, L.rptr.p, dx.p, dy.p, N, bnum), "SptrsvCsr_l");
    dy.to_host(y);
}
// senk_sparse.hpp:359-379
template <typename T>
inline void SptrsvCsr_u(T *val, int *cind, int *rptr, T *x, T *y, int N, int bnum) {
    using namespace gpu_detail;
    Csr<Val<T>> U(val, cind, rptr, N);
    Dev<Val<T>> dx(x, N), dy(N);
    Plan p(rptr, cind, N, 1);
    check(trsv_bj(p.p, true, U.val.p, U.cind.p, U.rptr.p, dx.p, dy.p, N, bnum), "SptrsvCsr_u");
    dy.to_host(y);
}
// senk_sparse.hpp:392-416 (sequential ILUB forward solve)
template <typename T, int bnl, int bnw>
inline void SptrsvBcsr_l(T *bval, int *bcind, int *brptr, T *x, T *y, int N) {
    using namespace gpu_detail;
    Bcsr<Val<T>> L(bval, bcind, brptr, N, bnl, bnw);
    Dev<Val<T>> dx(x, N), dy(N);
    check(trsv_bcsr(false, L.val.p, L.cind.p, L.rptr.p, dx.p, dy.p, N, bnl, bnw), "SptrsvBcsr_l");
    dy.to_host(y);
}
// senk_sparse.hpp:429-465
template <typename T, int bnl, int bnw>
inline void SptrsvBcsr_u(T *bval, int *bcind, int *brptr, T *x, T *y, int N) {
    using namespace gpu_detail;
    Bcsr<Val<T>> U(bval, bcind, brptr, N, bnl, bnw);
    Dev<Val<T>> dx(x, N), dy(N);
    check(trsv_bcsr(true, U.val.p, U.cind.p, U.rptr.p, dx.p, dy.p, N, bnl, bnw), "SptrsvBcsr_u");
    dy.to_host(y);
}
// senk_sparse.hpp:481-517 (ABMC)
template <typename T, int bnl, int bnw>
inline void SptrsvBcsr_l(T *bval, int *bcind, int *brptr, T *x, T *y, int N, int *cptr, int cnum, int bsize) {
    using namespace gpu_detail;
    Bcsr<Val<T>> L(bval, bcind, brptr, N, bnl, bnw);
    Dev<Val<T>> dx(x, N), dy(N);
    check(trsv_bcsr_abmc(false, L.val.p, L.cind.p, L.rptr.p, dx.p, dy.p, N, cptr, cnum, bsize, bnl, bnw),
          "SptrsvBcsr_l");
    dy.to_host(y);
}
// senk_sparse.hpp:533-581
template <typename T, int bnl, int bnw>
inline void SptrsvBcsr_u(T *bval, int *bcind, int *brptr, T *x, T *y, int N, int *cptr, int cnum, int bsize) {
    using namespace gpu_detail;
    Bcsr<Val<T>> U(bval, bcind, brptr, N, bnl, bnw);
    Dev<Val<T>> dx(x, N), dy(N);
    check(trsv_bcsr_abmc(true, U.val.p, U.cind.p, U.rptr.p, dx.p, dy.p, N, cptr, cnum, bsize, bnl, bnw),
          "SptrsvBcsr_u");
    dy.to_host(y);
}

}  // namespace sparse

namespace blas1 {

// senk_blas1.hpp:26-30
template <typename T>
inline void Copy(T *x, T *y, int N) {
    using namespace gpu_detail;
    Dev<Val<T>> dx(x, N), dy(N);
    check(copy(dx.p, dy.p, N), "Copy");
    dy.to_host(y);
}
// senk_blas1.hpp:38-42
template <typename T>
inline void Scal(T a, T *x, int N) {
    using namespace gpu_detail;
    Dev<Val<T>> dx(x, N);
    check(scal((Val<T>)a, dx.p, N), "Scal");
    dx.to_host(x);
}
// senk_blas1.hpp:51-55
template <typename T>
inline void Axpy(T a, T *x, T *y, int N) {
    using namespace gpu_detail;
    Dev<Val<T>> dx(x, N), dy(y, N);
    check(axpy((Val<T>)a, dx.p, dy.p, N), "Axpy");
    dy.to_host(y);
}
// senk_blas1.hpp:65-69
template <typename T>
inline void Axpby(T a, T *x, T b, T *y, int N) {
    using namespace gpu_detail;
    Dev<Val<T>> dx(x, N), dy(y, N);
    check(axpby((Val<T>)a, dx.p, (Val<T>)b, dy.p, N), "Axpby");
    dy.to_host(y);
}
// senk_blas1.hpp:79-83
template <typename T>
inline void Axpyz(T a, T *x, T *y, T *z, int N) {
    using namespace gpu_detail;
    Dev<Val<T>> dx(x, N), dy(y, N), dz(N);
    check(axpyz((Val<T>)a, dx.p, dy.p, dz.p, N), "Axpyz");
    dz.to_host(z);
}
// senk_blas1.hpp:92-98
template <typename T>
inline T Dot(T *x, T *y, int N) {
    using namespace gpu_detail;
    Dev<Val<T>> dx(x, N), dy(y, N);
    double r = 0;
    check(dot(dx.p, dy.p, N, &r), "Dot");
    return (T)r;
}
// senk_blas1.hpp:106-112
template <typename T>
inline T Nrm2(T *x, int N) {
    using namespace gpu_detail;
    Dev<Val<T>> dx(x, N);
    double r = 0;
    check(dot(dx.p, dx.p, N, &r), "Nrm2");
    return (T)std::sqrt(r);
}
// senk_blas1.hpp:120-124
template <typename T>
inline void HadProd(T *x, T *y, int N) {
    using namespace gpu_detail;
    Dev<Val<T>> dx(x, N), dy(y, N);
    check(hadprod(dx.p, dy.p, N), "HadProd");
    dy.to_host(y);
}
// senk_blas1.hpp:132-136
template <typename T>
inline void HadDiv(T *x, T *y, int N) {
    using namespace gpu_detail;
    Dev<Val<T>> dx(x, N), dy(y, N);
    check(haddiv(dx.p, dy.p, N), "HadDiv");
    dy.to_host(y);
}
// senk_blas1.hpp:145-153 (r = sqrt(a^2 + b^2), c = a/r, s = -b/r; a scalar host routine)
template <typename T>
inline T Ggen(T a, T b, T *c, T *s) {
    double cc = 0, ss = 0;
    const double r = senk_ggen((double)a, (double)b, &cc, &ss);
    *c = (T)cc;
    *s = (T)ss;
    return (T)r;
}
// senk_blas1.hpp:162-168
template <typename T>
inline void Grot(T c, T s, T *a, T *b) {
    double aa = (double)*a, bb = (double)*b;
    senk_grot((double)c, (double)s, &aa, &bb);
    *a = (T)aa;
    *b = (T)bb;
}

}  // namespace blas1

namespace blas2 {

// senk_blas2.hpp:24-34 (column-major upper triangular U, leading dimension n, order m)
template <typename T>
inline void Trsv(T *U, T *b, T *x, int n, int m) {
    if constexpr (std::is_same<T, double>::value) {
        senk_trsv(U, b, x, n, m);
    } else {
        const size_t k = m > 0 ? (size_t)m : 0;
        std::vector<double> u(U, U + (size_t)n * k), bb(b, b + k), xx(k);
        senk_trsv(u.data(), bb.data(), xx.data(), n, m);
        for (size_t i = 0; i < k; i++) x[i] = (T)xx[i];
    }
}

}  // namespace blas2

namespace test {

// senk_test.hpp:11-22: prints "# test <||b - A x|| / ||b||>" (FP64 on the device)
inline void RelativeResidualError(double *val, int *cind, int *rptr, double *b, double *x, int N) {
    using namespace gpu_detail;
    Csr<double> A(val, cind, rptr, N);
    Dev<double> db(b, N), dx(x, N);
    double r = 0;
    check(senk_relative_residual_error(ctx(), A.val.p, A.cind.p, A.rptr.p, db.p, dx.p, N, 1, &r),
          "RelativeResidualError");
}

}  // namespace test
}  // namespace senk
