// Instantiates every template of the drop-in header for T = double and T = float, with
// the reference's own argument types, so that a syntax-only compile checks each overload
// a reference caller can reach.  When SENK_WITH_REFERENCE is defined the reference's
// top-level header is included after the drop-in, as a caller keeping the reference's
// host setup code would do.
#include "senk_gpu.hpp"
#ifdef SENK_WITH_REFERENCE
#include "senk.hpp"
#endif

template <typename T>
void all(T *v, int *c, int *r, T *x, T *y, T *z, int *cptr, T a, int N) {
    using namespace senk;
    solver::Gmresm<T>(v, c, r, x, y, a, 1, 2, N, a);
    solver::Gmresm<T, 3, 3>(v, c, r, x, y, a, 1, 2, N, a);
    solver::IluGmresm<T>(v, c, r, v, c, r, v, c, r, x, y, a, 1, 2, N, a);
    solver::AmcIluGmresm<T>(v, c, r, v, c, r, v, c, r, cptr, 2, x, y, a, 1, 2, N, a);
    solver::AbmcIluGmresm<T>(v, c, r, v, c, r, v, c, r, cptr, 2, 4, x, y, a, 1, 2, N, a);
    solver::BjIluGmresm<T>(v, c, r, v, c, r, v, c, r, 4, x, y, a, 1, 2, N, a);
    solver::IlubGmresm<T, 8, 1>(v, c, r, v, c, r, v, c, r, x, y, a, 1, 2, N, a);
    solver::AbmcIlubGmresm<T, 3, 3>(v, c, r, v, c, r, v, c, r, cptr, 2, 6, x, y, a, 1, 2, N, a);
    solver::Gcrm<T>(v, c, r, x, y, a, 1, 2, N, a);
    solver::IluGcrm<T>(v, c, r, v, c, r, v, c, r, x, y, a, 1, 2, N, a);
    solver::IlubGcrm<T, 2, 2>(v, c, r, v, c, r, v, c, r, x, y, a, 1, 2, N, a);
    solver::Bicgstab<T>(v, c, r, x, y, a, 10, N, a);
    solver::IluBicgstab<T>(v, c, r, v, c, r, v, c, r, x, y, a, 10, N, a);
    solver::IlubBicgstab<T, 4, 2>(v, c, r, v, c, r, v, c, r, x, y, a, 10, N, a);
    sparse::SpmvCsr<T>(v, c, r, x, y, N);
    sparse::SpmvCsr<T>(v, c, r, z, x, y, N);
    sparse::SpmvBcsr<T, 3, 3>(v, c, r, x, y, N);
    sparse::SpmvSell<T>(v, c, r, 64, x, y, N);
    sparse::SptrsvCsr_l<T>(v, c, r, x, y, N);
    sparse::SptrsvCsr_u<T>(v, c, r, x, y, N);
    sparse::SptrsvCsr_l<T>(v, c, r, x, y, N, cptr, 3);
    sparse::SptrsvCsr_u<T>(v, c, r, x, y, N, cptr, 3);
    sparse::SptrsvCsr_l<T>(v, c, r, x, y, N, cptr, 3, 4);
    sparse::SptrsvCsr_u<T>(v, c, r, x, y, N, cptr, 3, 4);
    sparse::SptrsvCsr_l<T>(v, c, r, x, y, N, 4);
    sparse::SptrsvCsr_u<T>(v, c, r, x, y, N, 4);
    sparse::SptrsvBcsr_l<T, 3, 3>(v, c, r, x, y, N);
    sparse::SptrsvBcsr_u<T, 3, 3>(v, c, r, x, y, N);
    sparse::SptrsvBcsr_l<T, 8, 1>(v, c, r, x, y, N, cptr, 3, 8);
    sparse::SptrsvBcsr_u<T, 8, 1>(v, c, r, x, y, N, cptr, 3, 8);
    blas1::Copy<T>(x, y, N);
    blas1::Scal<T>(a, x, N);
    blas1::Axpy<T>(a, x, y, N);
    blas1::Axpby<T>(a, x, a, y, N);
    blas1::Axpyz<T>(a, x, y, z, N);
    T d = blas1::Dot<T>(x, y, N) + blas1::Nrm2<T>(x, N);
    blas1::HadProd<T>(x, y, N);
    blas1::HadDiv<T>(x, y, N);
    T cc, ss;
    d += blas1::Ggen<T>(a, d, &cc, &ss);
    blas1::Grot<T>(cc, ss, x, y);
    blas2::Trsv<T>(v, x, y, 31, 30);
}

int main(int argc, char **) {
    if (argc > 1000) {  // never runs: the point is the instantiation
        all<double>(nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 1.0, 0);
        all<float>(nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 1.0f, 0);
        senk::test::RelativeResidualError(nullptr, nullptr, nullptr, nullptr, nullptr, 0);
    }
    return 0;
}
