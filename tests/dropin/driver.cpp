// A program written against the reference's API (senk::solver::AbmcIluGmresm with host
// arrays), built against the drop-in header: it reads a reordered system and its
// split ILU(0) factors from a binary file, solves, and writes x.
#include <cstdio>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "senk_gpu.hpp"

template <typename T>
static std::vector<T> rd(FILE *f) {
    long long n = 0;
    if (std::fread(&n, sizeof n, 1, f) != 1) std::exit(2);
    std::vector<T> v((size_t)n);
    if (n && std::fread(v.data(), sizeof(T), (size_t)n, f) != (size_t)n) std::exit(2);
    return v;
}

int main(int argc, char **argv) {
    if (argc != 3) return 2;
    FILE *f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    auto hdr = rd<int>(f);  // N, cnum, bsize, outer, m
    auto val = rd<double>(f);
    auto cind = rd<int>(f);
    auto rptr = rd<int>(f);
    auto lval = rd<double>(f);
    auto lcind = rd<int>(f);
    auto lrptr = rd<int>(f);
    auto uval = rd<double>(f);
    auto ucind = rd<int>(f);
    auto urptr = rd<int>(f);
    auto cptr = rd<int>(f);
    auto b = rd<double>(f);
    auto eps = rd<double>(f);
    std::fclose(f);
    const int N = hdr[0], cnum = hdr[1], bsize = hdr[2], outer = hdr[3], m = hdr[4];
    double nrm_b = 0;
    for (double v : b) nrm_b += v * v;
    nrm_b = std::sqrt(nrm_b);
    std::vector<double> x((size_t)N, 0.0);
    senk::solver::AbmcIluGmresm<double>(val.data(), cind.data(), rptr.data(), lval.data(), lcind.data(),
                                        lrptr.data(), uval.data(), ucind.data(), urptr.data(), cptr.data(), cnum,
                                        bsize, b.data(), x.data(), nrm_b, outer, m, N, eps[0]);
    FILE *o = std::fopen(argv[2], "wb");
    if (!o) return 2;
    std::fwrite(x.data(), sizeof(double), (size_t)N, o);
    std::fclose(o);
    return 0;
}
