// Synthetic benchmark systems.  Every modelling choice is stated here once; the
// matrices are deterministic functions of their arguments (std::mt19937 streams from
// the seed), sorted CSR with a diagonal in every row.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "senk_amd.h"

namespace {

template <typename T>
T *to_malloc(const std::vector<T> &v) {
    T *p = (T *)std::malloc(sizeof(T) * (v.size() ? v.size() : 1));
    if (p && !v.empty()) std::memcpy(p, v.data(), sizeof(T) * v.size());
    return p;
}
int emit(const std::vector<int> &rp, const std::vector<int> &ci, const std::vector<double> &v, int n, int *N,
         int **rptr, int **cind, double **val) {
    *N = n;
    *rptr = to_malloc(rp);
    *cind = to_malloc(ci);
    *val = to_malloc(v);
    return (*rptr && *cind && *val) ? SENK_OK : SENK_ERR_ALLOC;
}

}  // namespace

extern "C" {

// 3D 7-point Poisson on an nx*ny*nz grid, Dirichlet boundary eliminated: diagonal 6,
// neighbours -1; unknown (x,y,z) -> x + nx*y + nx*ny*z.
int senk_gen_poisson7(int nx, int ny, int nz, int *N, int **rptr, int **cind, double **val) {
    if (nx <= 0 || ny <= 0 || nz <= 0 || !N || !rptr || !cind || !val) return SENK_ERR_ARG;
    const long long n = (long long)nx * ny * nz;
    if (n > 0x7fffffffLL / 8) return SENK_ERR_ARG;
    std::vector<int> rp((size_t)n + 1, 0), ci;
    std::vector<double> v;
    ci.reserve((size_t)n * 7);
    v.reserve((size_t)n * 7);
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                const int i = x + nx * (y + ny * z);
                auto add = [&](int c, double a) { ci.push_back(c); v.push_back(a); };
                if (z > 0) add(i - nx * ny, -1);
                if (y > 0) add(i - nx, -1);
                if (x > 0) add(i - 1, -1);
                add(i, 6);
                if (x < nx - 1) add(i + 1, -1);
                if (y < ny - 1) add(i + nx, -1);
                if (z < nz - 1) add(i + nx * ny, -1);
                rp[i + 1] = (int)ci.size();
            }
    return emit(rp, ci, v, (int)n, N, rptr, cind, val);
}

// 3D 27-point convection-diffusion on an n^3 grid of the unit cube, h = 1/(n+1),
// Dirichlet boundary eliminated, scaled by h^2:
//   diffusion: the 27-point Laplacian with centre 26 and all 26 neighbours -1, divided
//              by 9 (its second moment is 9 h^2 per axis, so L27/9 ~ -h^2 Laplacian);
//   convection: beta = peclet * (1, 0.5, 0.25) with central differences on the six
//              face neighbours, +-beta_d * h / 2.
// The pattern is symmetric, the values are not.  nnz = (3n-2)^3.
int senk_gen_convdiff27(int n, double peclet, int *N, int **rptr, int **cind, double **val) {
    if (n <= 0 || !N || !rptr || !cind || !val) return SENK_ERR_ARG;
    const long long nn = (long long)n * n * n;
    if (nn * 27 > 0x7fffffffLL) return SENK_ERR_ARG;
    const double h = 1.0 / (n + 1), beta[3] = {peclet * 1.0, peclet * 0.5, peclet * 0.25};
    std::vector<int> rp((size_t)nn + 1, 0), ci;
    std::vector<double> v;
    ci.reserve((size_t)nn * 27);
    v.reserve((size_t)nn * 27);
    for (int z = 0; z < n; z++)
        for (int y = 0; y < n; y++)
            for (int x = 0; x < n; x++) {
                const long long i = x + (long long)n * (y + (long long)n * z);
                for (int dz = -1; dz <= 1; dz++)
                    for (int dy = -1; dy <= 1; dy++)
                        for (int dx = -1; dx <= 1; dx++) {
                            const int X = x + dx, Y = y + dy, Z = z + dz;
                            if (X < 0 || Y < 0 || Z < 0 || X >= n || Y >= n || Z >= n) continue;
                            const int nzc = (dx != 0) + (dy != 0) + (dz != 0);
                            double a = nzc == 0 ? 26.0 / 9.0 : -1.0 / 9.0;
                            if (nzc == 1) {
                                const int d = dx ? 0 : dy ? 1 : 2, s = dx + dy + dz;
                                a += s * beta[d] * h / 2;
                            }
                            ci.push_back((int)(X + (long long)n * (Y + (long long)n * Z)));
                            v.push_back(a);
                        }
                rp[i + 1] = (int)ci.size();
            }
    return emit(rp, ci, v, (int)nn, N, rptr, cind, val);
}

// 3D 7-point variable-coefficient diffusion -div(k grad u) on an n^3 grid of cells.
// log k is a Gaussian random field: white noise N(0,1) per cell (mt19937(seed),
// std::normal_distribution, x fastest), smoothed by a separable periodic Gaussian
// filter of standard deviation corr_len/2 cells (truncated at 3 standard deviations),
// then shifted and scaled to zero mean and standard deviation sigma.  Face
// coefficients are harmonic means of the two cells; a Dirichlet boundary face uses the
// cell's own k.  Row i: -k_f for each neighbour, diagonal = sum of its faces' k_f.
int senk_gen_vardiff7(int n, int seed, double corr_len, double sigma, int *N, int **rptr, int **cind, double **val) {
    if (n <= 0 || corr_len <= 0 || !N || !rptr || !cind || !val) return SENK_ERR_ARG;
    const size_t nn = (size_t)n * n * n;
    std::vector<double> g(nn), t(nn);
    std::mt19937 eng((unsigned)seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    for (size_t i = 0; i < nn; i++) g[i] = nd(eng);
    const double s = corr_len / 2.0;
    const int R = std::max(1, (int)std::ceil(3 * s));
    std::vector<double> w((size_t)2 * R + 1);
    double ws = 0;
    for (int k = -R; k <= R; k++) ws += (w[k + R] = std::exp(-0.5 * k * k / (s * s)));
    for (auto &x : w) x /= ws;
    auto idx = [n](int x, int y, int z) { return (size_t)x + (size_t)n * ((size_t)y + (size_t)n * z); };
    for (int axis = 0; axis < 3; axis++) {
        for (int z = 0; z < n; z++)
            for (int y = 0; y < n; y++)
                for (int x = 0; x < n; x++) {
                    double acc = 0;
                    for (int k = -R; k <= R; k++) {
                        int X = x, Y = y, Z = z;
                        int &c = axis == 0 ? X : axis == 1 ? Y : Z;
                        c = ((c + k) % n + n) % n;
                        acc += w[k + R] * g[idx(X, Y, Z)];
                    }
                    t[idx(x, y, z)] = acc;
                }
        g.swap(t);
    }
    double mean = 0, var = 0;
    for (double x : g) mean += x;
    mean /= (double)nn;
    for (double x : g) var += (x - mean) * (x - mean);
    const double sd = std::sqrt(var / (double)nn);
    std::vector<double> k(nn);
    for (size_t i = 0; i < nn; i++) k[i] = std::exp(sigma * (g[i] - mean) / (sd > 0 ? sd : 1));
    std::vector<int> rp(nn + 1, 0), ci;
    std::vector<double> v;
    ci.reserve(nn * 7);
    v.reserve(nn * 7);
    auto hm = [](double a, double b) { return 2 * a * b / (a + b); };
    for (int z = 0; z < n; z++)
        for (int y = 0; y < n; y++)
            for (int x = 0; x < n; x++) {
                const size_t i = idx(x, y, z);
                const double ki = k[i];
                double diag = 0;
                const int off[6][3] = {{0, 0, -1}, {0, -1, 0}, {-1, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
                double kf[6];
                long long col[6];
                for (int f = 0; f < 6; f++) {
                    const int X = x + off[f][0], Y = y + off[f][1], Z = z + off[f][2];
                    if (X < 0 || Y < 0 || Z < 0 || X >= n || Y >= n || Z >= n) {
                        kf[f] = ki;
                        col[f] = -1;
                    } else {
                        kf[f] = hm(ki, k[idx(X, Y, Z)]);
                        col[f] = (long long)idx(X, Y, Z);
                    }
                    diag += kf[f];
                }
                for (int f = 0; f < 3; f++)
                    if (col[f] >= 0) { ci.push_back((int)col[f]); v.push_back(-kf[f]); }
                ci.push_back((int)i);
                v.push_back(diag);
                for (int f = 3; f < 6; f++)
                    if (col[f] >= 0) { ci.push_back((int)col[f]); v.push_back(-kf[f]); }
                rp[i + 1] = (int)ci.size();
            }
    return emit(rp, ci, v, (int)nn, N, rptr, cind, val);
}

// Unstructured nonsymmetric matrix: row i draws round(lognormal(ln 12, 0.6)) clipped to
// [3, 128] off-diagonal candidates with columns (i + round(N(0, 0.001 n))) mod n;
// duplicates and the row's own column are dropped.  Off-diagonal values N(0,1) rounded
// to FP32; the diagonal is 1.1 x the row's absolute off-diagonal sum, rounded to FP32
// (1 for a row left without off-diagonals, which only small n produces), so every
// value is exact in FP32.  One mt19937(seed) stream, rows in order.
int senk_gen_unstructured(int n, int seed, int *N, int **rptr, int **cind, double **val) {
    if (n <= 1 || !N || !rptr || !cind || !val) return SENK_ERR_ARG;
    std::mt19937 eng((unsigned)seed);
    std::lognormal_distribution<double> len_d(std::log(12.0), 0.6);
    std::normal_distribution<double> off_d(0.0, 0.001 * n);
    std::normal_distribution<double> val_d(0.0, 1.0);
    std::vector<int> rp((size_t)n + 1, 0), ci;
    std::vector<double> v;
    ci.reserve((size_t)n * 16);
    v.reserve((size_t)n * 16);
    std::vector<std::pair<int, double>> row;
    for (int i = 0; i < n; i++) {
        long long cnt = std::llround(len_d(eng));
        cnt = std::min<long long>(128, std::max<long long>(3, cnt));
        row.clear();
        for (long long q = 0; q < cnt; q++) {
            long long c = ((long long)i + std::llround(off_d(eng))) % n;
            if (c < 0) c += n;
            const double a = (double)(float)val_d(eng);
            if (c == i) continue;
            row.push_back({(int)c, a});
        }
        std::stable_sort(row.begin(), row.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
        row.erase(std::unique(row.begin(), row.end(), [](const auto &a, const auto &b) { return a.first == b.first; }),
                  row.end());
        double s = 0;
        for (auto &e : row) s += std::fabs(e.second);
        const double d = s > 0 ? (double)(float)(1.1 * s) : 1.0;
        bool placed = false;
        for (auto &e : row) {
            if (!placed && e.first > i) { ci.push_back(i); v.push_back(d); placed = true; }
            ci.push_back(e.first);
            v.push_back(e.second);
        }
        if (!placed) { ci.push_back(i); v.push_back(d); }
        rp[i + 1] = (int)ci.size();
    }
    return emit(rp, ci, v, n, N, rptr, cind, val);
}

// Elasticity-like 3-dof system on an n^3 node grid (dof 3*node + component, node
// x-fastest).  Young's modulus per node E ~ lognormal(0, 1) (mt19937(seed)); a face
// between two nodes uses the harmonic mean of their E, a Dirichlet boundary face the
// node's own E.  With Lame parameters lambda = E nu / ((1+nu)(1-2nu)), mu = E/(2(1+nu)),
// the coupling of a face in axis d is the 3x3 block -(mu I + (lambda+mu) t_d t_d^T) with
// skewed unit directions t_x ~ (1, .2, .1), t_y ~ (.1, 1, .2), t_z ~ (.2, .1, 1) (so that
// every block is dense), and the diagonal block is the sum of the face blocks of all six
// faces.  The matrix is symmetric positive definite; stored as scalar CSR with dense
// 3x3 blocks (BCSR 3x3 recovers them exactly).
int senk_gen_elasticity3(int n, int seed, double nu, int *N, int **rptr, int **cind, double **val) {
    if (n <= 0 || nu <= -1 || nu >= 0.5 || !N || !rptr || !cind || !val) return SENK_ERR_ARG;
    const long long nodes = (long long)n * n * n;
    if (nodes * 3 * 21 > 0x7fffffffLL) return SENK_ERR_ARG;
    std::mt19937 eng((unsigned)seed);
    std::lognormal_distribution<double> ed(0.0, 1.0);
    std::vector<double> E((size_t)nodes);
    for (auto &e : E) e = ed(eng);
    double t[3][3] = {{1, .2, .1}, {.1, 1, .2}, {.2, .1, 1}};
    for (auto &r : t) {
        const double s = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        for (double &q : r) q /= s;
    }
    auto block = [&](double e, int d, double B[3][3]) {
        const double lam = e * nu / ((1 + nu) * (1 - 2 * nu)), mu = e / (2 * (1 + nu));
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) B[a][b] = (a == b ? mu : 0.0) + (lam + mu) * t[d][a] * t[d][b];
    };
    auto hm = [](double a, double b) { return 2 * a * b / (a + b); };
    auto nid = [n](int x, int y, int z) { return (long long)x + (long long)n * ((long long)y + (long long)n * z); };
    const int Nd = (int)(nodes * 3);
    std::vector<int> rp((size_t)Nd + 1, 0), ci;
    std::vector<double> v;
    ci.reserve((size_t)Nd * 21);
    v.reserve((size_t)Nd * 21);
    const int off[6][3] = {{0, 0, -1}, {0, -1, 0}, {-1, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int z = 0; z < n; z++)
        for (int y = 0; y < n; y++)
            for (int x = 0; x < n; x++) {
                const long long i = nid(x, y, z);
                double D[3][3] = {{0}}, Bf[6][3][3];
                long long col[6];
                for (int f = 0; f < 6; f++) {
                    const int X = x + off[f][0], Y = y + off[f][1], Z = z + off[f][2];
                    const int d = off[f][0] ? 0 : off[f][1] ? 1 : 2;
                    const bool in = X >= 0 && Y >= 0 && Z >= 0 && X < n && Y < n && Z < n;
                    const double e = in ? hm(E[i], E[nid(X, Y, Z)]) : E[i];
                    block(e, d, Bf[f]);
                    col[f] = in ? nid(X, Y, Z) : -1;
                    for (int a = 0; a < 3; a++)
                        for (int b = 0; b < 3; b++) D[a][b] += Bf[f][a][b];
                }
                for (int a = 0; a < 3; a++) {
                    for (int f = 0; f < 3; f++)
                        if (col[f] >= 0)
                            for (int b = 0; b < 3; b++) { ci.push_back((int)(3 * col[f] + b)); v.push_back(-Bf[f][a][b]); }
                    for (int b = 0; b < 3; b++) { ci.push_back((int)(3 * i + b)); v.push_back(D[a][b]); }
                    for (int f = 3; f < 6; f++)
                        if (col[f] >= 0)
                            for (int b = 0; b < 3; b++) { ci.push_back((int)(3 * col[f] + b)); v.push_back(-Bf[f][a][b]); }
                    rp[3 * i + a + 1] = (int)ci.size();
                }
            }
    return emit(rp, ci, v, Nd, N, rptr, cind, val);
}

// x* ~ U[-1, 1]: std::mt19937(seed) with std::uniform_real_distribution<double>(-1, 1) in
// index order, as the reference's driver draws b (main.cpp:117-119).
int senk_gen_uniform(int seed, int n, double *out) {
    if (n < 0 || (!out && n)) return SENK_ERR_ARG;
    std::mt19937 eng((unsigned)seed);
    std::uniform_real_distribution<> d(-1.0, 1.0);
    for (int i = 0; i < n; i++) out[i] = d(eng);
    return SENK_OK;
}

}  // extern "C"
