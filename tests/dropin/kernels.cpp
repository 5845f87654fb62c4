// A program written against the reference's kernel API (senk::sparse, senk::blas1,
// senk::blas2, senk::test with host arrays), built against the drop-in header.  It reads
// named arrays from a file, runs every sparse kernel form and BLAS-1/2 routine once for
// T = double and once for T = float (inputs rounded to float), and writes each output as
// a named array "<op>_d" / "<op>_f".
//
// File format, in and out: records of [int32 name length][name][int32 kind: 0 int32,
// 1 float64, 2 float32][int64 count][count elements].
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <string>
#include <vector>

#include "senk_gpu.hpp"

namespace {
struct Arr {
    int kind = 0;
    std::vector<int> i;
    std::vector<double> d;
    std::vector<float> f;
};
std::map<std::string, Arr> in;
FILE *out = nullptr;

bool read_all(const char *path) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    for (;;) {
        int32_t len = 0;
        if (std::fread(&len, 4, 1, f) != 1) break;
        std::string name((size_t)len, '\0');
        int32_t kind = 0;
        int64_t n = 0;
        if (std::fread(&name[0], 1, (size_t)len, f) != (size_t)len || std::fread(&kind, 4, 1, f) != 1 ||
            std::fread(&n, 8, 1, f) != 1)
            return false;
        Arr a;
        a.kind = kind;
        size_t got = 0;
        if (kind == 0) {
            a.i.resize((size_t)n);
            got = std::fread(a.i.data(), 4, (size_t)n, f);
        } else {
            a.d.resize((size_t)n);
            got = std::fread(a.d.data(), 8, (size_t)n, f);
        }
        if (got != (size_t)n) return false;
        in[name] = a;
    }
    std::fclose(f);
    return true;
}
int *I(const char *name) {
    auto it = in.find(name);
    if (it == in.end() || it->second.kind != 0) {
        std::fprintf(stderr, "missing int array %s\n", name);
        std::exit(3);
    }
    return it->second.i.data();
}
int I0(const char *name) { return I(name)[0]; }
template <typename T>
std::vector<T> V(const char *name) {
    auto it = in.find(name);
    if (it == in.end() || it->second.kind != 1) {
        std::fprintf(stderr, "missing real array %s\n", name);
        std::exit(3);
    }
    return std::vector<T>(it->second.d.begin(), it->second.d.end());
}
template <typename T>
void put(const std::string &name, const T *p, size_t n) {
    const int32_t len = (int32_t)name.size(), kind = sizeof(T) == 8 ? 1 : 2;
    const int64_t cnt = (int64_t)n;
    std::fwrite(&len, 4, 1, out);
    std::fwrite(name.data(), 1, name.size(), out);
    std::fwrite(&kind, 4, 1, out);
    std::fwrite(&cnt, 8, 1, out);
    std::fwrite(p, sizeof(T), n, out);
}

template <typename T>
void run(const char *suf) {
    using namespace senk;
    const int N = I0("N");
    const std::string s(suf);
    auto x = V<T>("x"), y0 = V<T>("y0"), sc = V<T>("scalars");
    std::vector<T> y((size_t)N);
    const T a = sc[0], b = sc[1];

    auto val = V<T>("A_val");
    sparse::SpmvCsr<T>(val.data(), I("A_cind"), I("A_rptr"), x.data(), y.data(), N);
    put("spmv" + s, y.data(), y.size());
    auto oval = V<T>("Aoff_val"), diag = V<T>("diag");
    sparse::SpmvCsr<T>(oval.data(), I("Aoff_cind"), I("Aoff_rptr"), diag.data(), x.data(), y.data(), N);
    put("spmv_diag" + s, y.data(), y.size());
    auto bval = V<T>("Ab_val");
    sparse::SpmvBcsr<T, 3, 3>(bval.data(), I("Ab_cind"), I("Ab_rptr"), x.data(), y.data(), N);
    put("spmv_bcsr" + s, y.data(), y.size());
    auto sval = V<T>("As_val");
    sparse::SpmvSell<T>(sval.data(), I("As_cind"), I("As_wid"), I0("As_len"), x.data(), y.data(), N);
    put("spmv_sell" + s, y.data(), y.size());

    auto tri = [&](const char *tag, const char *lname, const char *uname, auto &&lsolve, auto &&usolve) {
        std::string ln(lname), un(uname);
        auto lv = V<T>((ln + "_val").c_str()), uv = V<T>((un + "_val").c_str());
        std::vector<T> t((size_t)N), u((size_t)N);
        lsolve(lv.data(), I((ln + "_cind").c_str()), I((ln + "_rptr").c_str()), x.data(), t.data());
        usolve(uv.data(), I((un + "_cind").c_str()), I((un + "_rptr").c_str()), t.data(), u.data());
        put(std::string(tag) + "_l" + s, t.data(), t.size());
        put(std::string(tag) + "_u" + s, u.data(), u.size());
    };
    tri("seq", "L", "U", [&](T *v, int *c, int *r, T *xx, T *yy) { sparse::SptrsvCsr_l<T>(v, c, r, xx, yy, N); },
        [&](T *v, int *c, int *r, T *xx, T *yy) { sparse::SptrsvCsr_u<T>(v, c, r, xx, yy, N); });
    int *cm = I("amc_cptr"), *ca = I("abmc_cptr");
    const int nm = I0("amc_cnum"), na = I0("abmc_cnum"), bs = I0("abmc_bsize"), bnum = I0("bj_bnum");
    tri("amc", "Lm", "Um", [&](T *v, int *c, int *r, T *xx, T *yy) { sparse::SptrsvCsr_l<T>(v, c, r, xx, yy, N, cm, nm); },
        [&](T *v, int *c, int *r, T *xx, T *yy) { sparse::SptrsvCsr_u<T>(v, c, r, xx, yy, N, cm, nm); });
    tri("abmc", "La", "Ua",
        [&](T *v, int *c, int *r, T *xx, T *yy) { sparse::SptrsvCsr_l<T>(v, c, r, xx, yy, N, ca, na, bs); },
        [&](T *v, int *c, int *r, T *xx, T *yy) { sparse::SptrsvCsr_u<T>(v, c, r, xx, yy, N, ca, na, bs); });
    tri("bj", "Lj", "Uj", [&](T *v, int *c, int *r, T *xx, T *yy) { sparse::SptrsvCsr_l<T>(v, c, r, xx, yy, N, bnum); },
        [&](T *v, int *c, int *r, T *xx, T *yy) { sparse::SptrsvCsr_u<T>(v, c, r, xx, yy, N, bnum); });
    tri("bcsr", "Lb", "Ub",
        [&](T *v, int *c, int *r, T *xx, T *yy) { sparse::SptrsvBcsr_l<T, 3, 3>(v, c, r, xx, yy, N); },
        [&](T *v, int *c, int *r, T *xx, T *yy) { sparse::SptrsvBcsr_u<T, 3, 3>(v, c, r, xx, yy, N); });
    int *cb = I("bcsra_cptr");
    const int nb = I0("bcsra_cnum"), bsb = I0("bcsra_bsize");
    tri("bcsra", "Lba", "Uba",
        [&](T *v, int *c, int *r, T *xx, T *yy) { sparse::SptrsvBcsr_l<T, 3, 3>(v, c, r, xx, yy, N, cb, nb, bsb); },
        [&](T *v, int *c, int *r, T *xx, T *yy) { sparse::SptrsvBcsr_u<T, 3, 3>(v, c, r, xx, yy, N, cb, nb, bsb); });

    std::vector<T> z((size_t)N), w = y0;
    blas1::Copy<T>(x.data(), z.data(), N);
    put("copy" + s, z.data(), z.size());
    z = x;
    blas1::Scal<T>(a, z.data(), N);
    put("scal" + s, z.data(), z.size());
    w = y0;
    blas1::Axpy<T>(a, x.data(), w.data(), N);
    put("axpy" + s, w.data(), w.size());
    w = y0;
    blas1::Axpby<T>(a, x.data(), b, w.data(), N);
    put("axpby" + s, w.data(), w.size());
    blas1::Axpyz<T>(a, x.data(), y0.data(), z.data(), N);
    put("axpyz" + s, z.data(), z.size());
    w = y0;
    blas1::HadProd<T>(x.data(), w.data(), N);
    put("hadprod" + s, w.data(), w.size());
    w = y0;
    blas1::HadDiv<T>(x.data(), w.data(), N);
    put("haddiv" + s, w.data(), w.size());
    T red[2] = {blas1::Dot<T>(x.data(), y0.data(), N), blas1::Nrm2<T>(x.data(), N)};
    put("dot_nrm2" + s, red, 2);
    T c = 0, sn = 0;
    T g[5];
    g[0] = blas1::Ggen<T>(a, b, &c, &sn);
    g[1] = c;
    g[2] = sn;
    g[3] = a;
    g[4] = b;
    blas1::Grot<T>(c, sn, &g[3], &g[4]);
    put("ggen_grot" + s, g, 5);
    auto H = V<T>("H"), e = V<T>("e");
    const int n = I0("trsv_n"), m = I0("trsv_m");
    std::vector<T> yt((size_t)m);
    blas2::Trsv<T>(H.data(), e.data(), yt.data(), n, m);
    put("trsv" + s, yt.data(), yt.size());
}
}  // namespace

int main(int argc, char **argv) {
    if (argc != 3 || !read_all(argv[1])) return 2;
    out = std::fopen(argv[2], "wb");
    if (!out) return 2;
    run<double>("_d");
    run<float>("_f");
    std::fclose(out);
    auto val = V<double>("A_val"), x = V<double>("x"), y0 = V<double>("y0");
    // prints "# test <relres>" for b = y0 and the solution guess x
    senk::test::RelativeResidualError(val.data(), I("A_cind"), I("A_rptr"), y0.data(), x.data(), I0("N"));
    return 0;
}
