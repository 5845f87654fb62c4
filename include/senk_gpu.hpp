// C++ drop-in for SenK's preconditioned Krylov drivers (senk_gmres.hpp, senk_gcr.hpp,
// senk_bicgstab.hpp) and the SpMV / colored SpTRSV kernels (senk_sparse.hpp): the same
// namespaces, template names, argument order and meaning, running on the GPU through
// the C ABI of senk_amd.h.  Replace `#include "senk.hpp"` by `#include "senk_gpu.hpp"`
// for these calls and link -lsenk_amd.
//
// Host-pointer mode: every call uploads its arrays, runs on the device and copies the
// outputs back (x of a solver, y of a kernel), so a single call pays the PCIe transfers;
// the handle API of senk_amd.h keeps data resident for repeated solves.  The solvers
// print the lines the reference prints.  Errors print the library's message and call
// exit(1), as the reference does.  float instantiations compute in double and round the
// outputs back to float.  One device context per thread, created on first use.
#pragma once

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
// A device copy of a host array (converted to double for floating point).
template <typename D>
struct Dev {
    D *p = nullptr;
    size_t n = 0;
    template <typename H>
    Dev(const H *h, size_t count) : n(count) {
        check(senk_device_malloc(ctx(), sizeof(D) * (n ? n : 1), (void **)&p), "senk_device_malloc");
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
template <typename T>
struct Csr {
    Dev<double> val;
    Dev<int> cind, rptr;
    Csr(const T *v, const int *c, const int *r, int N) : val(v, (size_t)r[N]), cind(c, (size_t)r[N]), rptr(r, (size_t)N + 1) {}
};
template <typename T>
struct Bcsr {
    Dev<double> val;
    Dev<int> cind, rptr;
    Bcsr(const T *v, const int *c, const int *r, int N, int bnl, int bnw)
        : val(v, (size_t)r[N / bnl] * bnl * bnw), cind(c, (size_t)r[N / bnl]), rptr(r, (size_t)N / bnl + 1) {}
};
}  // namespace gpu_detail

namespace solver {

// senk_gmres.hpp:33-95
template <typename T>
void Gmresm(T *val, int *cind, int *rptr, T *b, T *x, T nrm_b, int outer, int m, int N, T epsilon) {
    using namespace gpu_detail;
    Csr<T> A(val, cind, rptr, N);
    Dev<double> db(b, N), dx(x, N);
    check(senk_gmresm_f64(ctx(), A.val.p, A.cind.p, A.rptr.p, db.p, dx.p, nrm_b, outer, m, N, epsilon, 1, nullptr,
                          nullptr, 0), "Gmresm");
    dx.to_host(x);
}
// senk_gmres.hpp:113-175 (BCSR matrix)
template <typename T, int bnl, int bnw>
void Gmresm(T *bval, int *bcind, int *brptr, T *b, T *x, T nrm_b, int outer, int m, int N, T epsilon) {
    using namespace gpu_detail;
    Bcsr<T> A(bval, bcind, brptr, N, bnl, bnw);
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
    Csr<T> A(val, cind, rptr, N), L(lval, lcind, lrptr, N), U(uval, ucind, urptr, N);
    Dev<double> db(b, N), dx(x, N);
    senk_trsv_plan lp, up;
    check(senk_trsv_plan_create(ctx(), lrptr, lcind, N, 0, &lp), "IluGmresm");
    check(senk_trsv_plan_create(ctx(), urptr, ucind, N, 1, &up), "IluGmresm");
    check(senk_ilu_gmresm_f64(ctx(), A.val.p, A.cind.p, A.rptr.p, L.val.p, L.cind.p, L.rptr.p, U.val.p, U.cind.p,
                              U.rptr.p, lp, up, db.p, dx.p, nrm_b, outer, m, N, epsilon, 1, nullptr, nullptr, 0),
          "IluGmresm");
    senk_trsv_plan_destroy(lp);
    senk_trsv_plan_destroy(up);
    dx.to_host(x);
}
// senk_gmres.hpp:293-367
template <typename T>
void AmcIluGmresm(T *val, int *cind, int *rptr, T *lval, int *lcind, int *lrptr, T *uval, int *ucind, int *urptr,
                  int *cptr, int cnum, T *b, T *x, T nrm_b, int outer, int m, int N, T epsilon) {
    using namespace gpu_detail;
    Csr<T> A(val, cind, rptr, N), L(lval, lcind, lrptr, N), U(uval, ucind, urptr, N);
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
    Csr<T> A(val, cind, rptr, N), L(lval, lcind, lrptr, N), U(uval, ucind, urptr, N);
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
    Csr<T> A(val, cind, rptr, N), L(lval, lcind, lrptr, N), U(uval, ucind, urptr, N);
    Dev<double> db(b, N), dx(x, N);
    senk_trsv_plan lp, up;
    check(senk_trsv_plan_create(ctx(), lrptr, lcind, N, 0, &lp), "BjIluGmresm");
    check(senk_trsv_plan_create(ctx(), urptr, ucind, N, 1, &up), "BjIluGmresm");
    check(senk_bj_ilu_gmresm_f64(ctx(), A.val.p, A.cind.p, A.rptr.p, L.val.p, L.cind.p, L.rptr.p, U.val.p, U.cind.p,
                                 U.rptr.p, lp, up, bnum, db.p, dx.p, nrm_b, outer, m, N, epsilon, 1, nullptr, nullptr,
                                 0), "BjIluGmresm");
    senk_trsv_plan_destroy(lp);
    senk_trsv_plan_destroy(up);
    dx.to_host(x);
}
// senk_gmres.hpp:582-655
template <typename T, int bnl, int bnw>
void IlubGmresm(T *val, int *cind, int *rptr, T *blval, int *blcind, int *blrptr, T *buval, int *bucind,
                int *burptr, T *b, T *x, T nrm_b, int outer, int m, int N, T epsilon) {
    using namespace gpu_detail;
    Csr<T> A(val, cind, rptr, N);
    Bcsr<T> L(blval, blcind, blrptr, N, bnl, bnw), U(buval, bucind, burptr, N, bnl, bnw);
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
    Csr<T> A(val, cind, rptr, N);
    Bcsr<T> L(blval, blcind, blrptr, N, bnl, bnw), U(buval, bucind, burptr, N, bnl, bnw);
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
    Csr<T> A(val, cind, rptr, N);
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
    Csr<T> A(val, cind, rptr, N), L(lval, lcind, lrptr, N), U(uval, ucind, urptr, N);
    Dev<double> db(b, N), dx(x, N);
    senk_trsv_plan lp, up;
    check(senk_trsv_plan_create(ctx(), lrptr, lcind, N, 0, &lp), "IluGcrm");
    check(senk_trsv_plan_create(ctx(), urptr, ucind, N, 1, &up), "IluGcrm");
    check(senk_ilu_gcrm_f64(ctx(), A.val.p, A.cind.p, A.rptr.p, L.val.p, L.cind.p, L.rptr.p, U.val.p, U.cind.p,
                            U.rptr.p, lp, up, db.p, dx.p, nrm_b, outer, m, N, epsilon, 1, nullptr, nullptr, 0),
          "IluGcrm");
    senk_trsv_plan_destroy(lp);
    senk_trsv_plan_destroy(up);
    dx.to_host(x);
}
// senk_gcr.hpp:163-212
template <typename T, int bnl, int bnw>
void IlubGcrm(T *val, int *cind, int *rptr, T *blval, int *blcind, int *blrptr, T *buval, int *bucind, int *burptr,
              T *b, T *x, T nrm_b, int outer, int m, int N, T epsilon) {
    using namespace gpu_detail;
    Csr<T> A(val, cind, rptr, N);
    Bcsr<T> L(blval, blcind, blrptr, N, bnl, bnw), U(buval, bucind, burptr, N, bnl, bnw);
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
    Csr<T> A(val, cind, rptr, N);
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
    Csr<T> A(val, cind, rptr, N), L(lval, lcind, lrptr, N), U(uval, ucind, urptr, N);
    Dev<double> db(b, N), dx(x, N);
    senk_trsv_plan lp, up;
    check(senk_trsv_plan_create(ctx(), lrptr, lcind, N, 0, &lp), "IluBicgstab");
    check(senk_trsv_plan_create(ctx(), urptr, ucind, N, 1, &up), "IluBicgstab");
    check(senk_ilu_bicgstab_f64(ctx(), A.val.p, A.cind.p, A.rptr.p, L.val.p, L.cind.p, L.rptr.p, U.val.p, U.cind.p,
                                U.rptr.p, lp, up, db.p, dx.p, nrm_b, max_iter, N, epsilon, 1, nullptr, nullptr, 0),
          "IluBicgstab");
    senk_trsv_plan_destroy(lp);
    senk_trsv_plan_destroy(up);
    dx.to_host(x);
}
// senk_bicgstab.hpp:198-267
template <typename T, int bnl, int bnw>
void IlubBicgstab(T *val, int *cind, int *rptr, T *blval, int *blcind, int *blrptr, T *buval, int *bucind,
                  int *burptr, T *b, T *x, T nrm_b, int max_iter, int N, T epsilon) {
    using namespace gpu_detail;
    Csr<T> A(val, cind, rptr, N);
    Bcsr<T> L(blval, blcind, blrptr, N, bnl, bnw), U(buval, bucind, burptr, N, bnl, bnw);
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
    Csr<T> A(val, cind, rptr, N);
    Dev<double> dx(x, N), dy(y, N);
    check(senk_spmv_csr_f64(ctx(), A.val.p, A.cind.p, A.rptr.p, dx.p, dy.p, N), "SpmvCsr");
    dy.to_host(y);
}
// senk_sparse.hpp:250-275
template <typename T>
inline void SptrsvCsr_l(T *val, int *cind, int *rptr, T *x, T *y, int N, int *cptr, int cnum, int bsize) {
    using namespace gpu_detail;
    Csr<T> L(val, cind, rptr, N);
    Dev<double> dx(x, N), dy(y, N);
    check(senk_sptrsv_csr_l_abmc_f64(ctx(), L.val.p, L.cind.p, L.rptr.p, dx.p, dy.p, N, cptr, cnum, bsize),
          "SptrsvCsr_l");
    dy.to_host(y);
}
// senk_sparse.hpp:289-315
template <typename T>
inline void SptrsvCsr_u(T *val, int *cind, int *rptr, T *x, T *y, int N, int *cptr, int cnum, int bsize) {
    using namespace gpu_detail;
    Csr<T> U(val, cind, rptr, N);
    Dev<double> dx(x, N), dy(y, N);
    check(senk_sptrsv_csr_u_abmc_f64(ctx(), U.val.p, U.cind.p, U.rptr.p, dx.p, dy.p, N, cptr, cnum, bsize),
          "SptrsvCsr_u");
    dy.to_host(y);
}

}  // namespace sparse
}  // namespace senk
