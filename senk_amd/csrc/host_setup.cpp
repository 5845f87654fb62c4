// Host-side integer setup for the device path: orderings (AMC/ABMC), format
// conversions, ILU(p)/ILUB symbolic patterns, dependency levels.  Integer outputs are
// bit-identical to the reference's; the algorithms are restated for speed at the
// benchmark sizes (linear-time coloring, linked-list symbolic fill) and the
// equivalence of each with the reference routine is stated at the routine.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>

#include "internal.hpp"

namespace senk_amd {
namespace host {

namespace {

template <typename T>
T *dup_malloc(const std::vector<T> &v) {
    T *p = (T *)std::malloc(sizeof(T) * (v.size() ? v.size() : 1));
    if (p && !v.empty()) std::memcpy(p, v.data(), sizeof(T) * v.size());
    return p;
}

// Transpose pattern: tr (rows of A^T) sorted ascending by construction.
void transpose_pattern(const int *rptr, const int *cind, int N, std::vector<int> &tp, std::vector<int> &tc) {
    tp.assign((size_t)N + 1, 0);
    for (int j = 0; j < rptr[N]; j++) tp[cind[j] + 1]++;
    for (int c = 0; c < N; c++) tp[c + 1] += tp[c];
    tc.resize((size_t)rptr[N]);
    std::vector<int> fill(tp.begin(), tp.end() - 1);
    for (int i = 0; i < N; i++)
        for (int j = rptr[i]; j < rptr[i + 1]; j++) tc[fill[cind[j]]++] = i;
}

// GetAdjacency (senk_graph.hpp:150-209): isSym keeps the pattern; otherwise the sorted
// union of the row patterns of A and A^T.
void adjacency(const int *rptr, const int *cind, int N, bool isSym, std::vector<int> &ar, std::vector<int> &ac) {
    if (isSym) {
        ar.assign(rptr, rptr + N + 1);
        ac.assign(cind, cind + rptr[N]);
        return;
    }
    std::vector<int> tp, tc;
    transpose_pattern(rptr, cind, N, tp, tc);
    ar.assign((size_t)N + 1, 0);
    ac.clear();
    ac.reserve((size_t)rptr[N] * 2);
    for (int i = 0; i < N; i++) {
        int a = rptr[i], b = tp[i];
        const int ea = rptr[i + 1], eb = tp[i + 1];
        while (a < ea || b < eb) {
            const int ca = a < ea ? cind[a] : N, cb = b < eb ? tc[b] : N;
            if (ca < cb) { ac.push_back(ca); a++; }
            else if (ca == cb) { ac.push_back(ca); a++; b++; }
            else { ac.push_back(cb); b++; }
        }
        ar[i + 1] = (int)ac.size();
    }
}

// Coloring (senk_graph.hpp:218-238).  The reference colors in rounds: in round c,
// scanning ascending, an uncolored node takes c unless a node colored c earlier in the
// same round listed it as a neighbour.  By induction on the node index that is exactly
// first-fit: color(i) = the smallest c such that no j < i with color(j) = c has i in
// its row.  Computed here in one pass over the in-neighbours (the transpose).
int first_fit(const int *rptr, const int *cind, int N, std::vector<int> &color) {
    std::vector<int> tp, tc;
    transpose_pattern(rptr, cind, N, tp, tc);
    color.assign((size_t)N, 0);
    std::vector<int> seen;  // seen[c] == i+1: color c taken by a lower in-neighbour of i
    int ncol = 0;
    for (int i = 0; i < N; i++) {
        for (int q = tp[i]; q < tp[i + 1]; q++) {
            const int j = tc[q];
            if (j >= i) break;  // in-neighbours are sorted ascending
            const int c = color[j];
            if ((int)seen.size() <= c) seen.resize((size_t)c + 1, 0);
            seen[c] = i + 1;
        }
        int c = 1;
        while (c < (int)seen.size() && seen[c] == i + 1) c++;
        color[i] = c;
        ncol = std::max(ncol, c);
    }
    return ncol;
}

}  // namespace

// GetAMCPermutation (senk_graph.hpp:250-277)
int amc(const int *cind, const int *rptr, int N, bool isSym, std::vector<int> &size_color, std::vector<int> &LP,
        std::vector<int> &RP) {
    std::vector<int> ar, ac, color;
    adjacency(rptr, cind, N, isSym, ar, ac);
    const int nc = first_fit(ar.data(), ac.data(), N, color);
    size_color.assign((size_t)nc + 1, 0);
    for (int i = 0; i < N; i++) size_color[color[i]]++;
    for (int c = 0; c < nc; c++) size_color[c + 1] += size_color[c];
    std::vector<int> pos(size_color.begin(), size_color.end() - 1);
    LP.assign((size_t)N, 0);
    RP.assign((size_t)N, 0);
    for (int i = 0; i < N; i++) LP[pos[color[i] - 1]++] = i;
    for (int i = 0; i < N; i++) RP[LP[i]] = i;
    return nc;
}

// GetABMCPermutation (senk_graph.hpp:291-334) with SimpleBlocking (:27-66) or
// ConnectedBlocking (:77-139).
int abmc(const int *cind, const int *rptr, int N, int bsize, bool isSym, bool connect, std::vector<int> &size_color,
         std::vector<int> &LP, std::vector<int> &RP) {
    if (bsize <= 0 || N % bsize) return -1;
    const int nb = N / bsize;
    std::vector<int> members((size_t)N), owner((size_t)N);
    if (!connect) {
        std::iota(members.begin(), members.end(), 0);
        for (int i = 0; i < N; i++) owner[i] = i / bsize;
    } else {
        // Breadth-first aggregation in visit order; a block closes at bsize members and a
        // partly filled block continues with the next unvisited seed.
        std::vector<char> seen((size_t)N, 0);
        std::vector<int> queue;
        queue.reserve(64);
        int count = 0, fill = 0;
        for (int sd = 0; sd < N; sd++) {
            if (seen[sd]) continue;
            members[count] = sd;
            owner[sd] = count / bsize;
            count++;
            fill++;
            seen[sd] = 1;
            if (fill == bsize) { fill = 0; continue; }
            queue.clear();
            queue.push_back(sd);
            size_t head = 0;
            bool full = false;
            while (!full && head < queue.size()) {
                const int q = queue[head++];
                for (int j = rptr[q]; j < rptr[q + 1]; j++) {
                    const int v = cind[j];
                    if (seen[v]) continue;
                    members[count] = v;
                    owner[v] = count / bsize;
                    count++;
                    fill++;
                    seen[v] = 1;
                    if (fill == bsize) { fill = 0; full = true; break; }
                    queue.push_back(v);
                }
            }
        }
        for (int b = 0; b < nb; b++) std::sort(members.begin() + (size_t)b * bsize, members.begin() + (size_t)(b + 1) * bsize);
    }
    // block graph: distinct owners of the columns of the members' rows, ascending
    std::vector<int> br((size_t)nb + 1, 0), bc;
    bc.reserve((size_t)nb * 8);
    std::vector<int> stamp((size_t)nb, -1), tmp;
    for (int b = 0; b < nb; b++) {
        tmp.clear();
        for (int m = 0; m < bsize; m++) {
            const int r = members[(size_t)b * bsize + m];
            for (int j = rptr[r]; j < rptr[r + 1]; j++) {
                const int o = owner[cind[j]];
                if (stamp[o] != b) { stamp[o] = b; tmp.push_back(o); }
            }
        }
        std::sort(tmp.begin(), tmp.end());
        bc.insert(bc.end(), tmp.begin(), tmp.end());
        br[b + 1] = (int)bc.size();
    }
    std::vector<int> ar, ac, color;
    adjacency(br.data(), bc.data(), nb, isSym, ar, ac);
    const int nc = first_fit(ar.data(), ac.data(), nb, color);
    std::vector<int> cnt((size_t)nc + 1, 0);
    for (int b = 0; b < nb; b++) cnt[color[b]]++;
    size_color.assign((size_t)nc + 1, 0);
    for (int c = 0; c < nc; c++) size_color[c + 1] = size_color[c] + cnt[c + 1];
    std::vector<int> pos(size_color.begin(), size_color.end() - 1);
    LP.assign((size_t)N, 0);
    RP.assign((size_t)N, 0);
    for (int b = 0; b < nb; b++) {
        const int slot = pos[color[b] - 1]++;
        std::memcpy(&LP[(size_t)slot * bsize], &members[(size_t)b * bsize], sizeof(int) * bsize);
    }
    for (int i = 0; i < N; i++) RP[LP[i]] = i;
    return nc;
}

// Dependency levels of blocks of bsize rows.  Lower: a block depends on the blocks of
// its rows' columns below the block; upper: above it.  Levels are numbered so that
// running them in increasing order (lower) or decreasing order (upper) respects every
// dependency, matching how colors are run.
int levels(const int *rptr, const int *cind, int N, int bsize, bool upper, std::vector<int> &lev_ptr,
           std::vector<int> &blk) {
    const int nb = N / bsize;
    std::vector<int> depth((size_t)nb, 0);
    int maxd = 0;
    if (!upper) {
        for (int b = 0; b < nb; b++) {
            int d = 0;
            for (int r = b * bsize; r < (b + 1) * bsize; r++)
                for (int j = rptr[r]; j < rptr[r + 1]; j++) {
                    const int ob = cind[j] / bsize;
                    if (ob < b) d = std::max(d, depth[ob] + 1);
                }
            depth[b] = d;
            maxd = std::max(maxd, d);
        }
    } else {
        for (int b = nb - 1; b >= 0; b--) {
            int d = 0;
            for (int r = b * bsize; r < (b + 1) * bsize; r++)
                for (int j = rptr[r]; j < rptr[r + 1]; j++) {
                    const int ob = cind[j] / bsize;
                    if (ob > b) d = std::max(d, depth[ob] + 1);
                }
            depth[b] = d;
            maxd = std::max(maxd, d);
        }
        for (int b = 0; b < nb; b++) depth[b] = maxd - depth[b];
    }
    const int nlev = nb ? maxd + 1 : 0;
    lev_ptr.assign((size_t)nlev + 1, 0);
    for (int b = 0; b < nb; b++) lev_ptr[depth[b] + 1]++;
    for (int l = 0; l < nlev; l++) lev_ptr[l + 1] += lev_ptr[l];
    std::vector<int> pos(lev_ptr.begin(), lev_ptr.end() - 1);
    blk.assign((size_t)nb, 0);
    for (int b = 0; b < nb; b++) blk[pos[depth[b]]++] = b;
    return nlev;
}

bool schedule_ok(const int *rptr, const int *cind, int N, const Schedule &s, const std::vector<int> &blk,
                 bool lower_part, bool upper_part) {
    const int bs = s.bsize, nb = N / bs;
    std::vector<int> lev((size_t)nb, -1);
    for (int l = 0; l < s.nlev; l++)
        for (int q = s.lev_ptr[l]; q < s.lev_ptr[l + 1]; q++) {
            const int b = blk.empty() ? q : blk[q];
            if (b < 0 || b >= nb) return false;
            lev[b] = l;
        }
    for (int b = 0; b < nb; b++)
        if (lev[b] < 0) return false;
    for (int r = 0; r < N; r++) {
        const int b = r / bs;
        for (int j = rptr[r]; j < rptr[r + 1]; j++) {
            const int ob = cind[j] / bs;
            if (ob == b) continue;
            if (cind[j] < r && lower_part && !(lev[ob] < lev[b])) return false;
            if (cind[j] > r && upper_part && !(lev[ob] > lev[b])) return false;
        }
    }
    return true;
}

}  // namespace host
}  // namespace senk_amd

using namespace senk_amd;
using senk_amd::host::dup_malloc;

// ===================================================================== C ABI: setup
extern "C" {

int senk_check_structure(const int *cind, const int *rptr, int N) {
    if (!cind || !rptr || N < 0) return 0;
    for (int i = 0; i < N; i++) {
        bool found = false;
        for (int j = rptr[i]; j < rptr[i + 1]; j++)
            if (cind[j] == i) { found = true; break; }
        if (!found) return 0;
    }
    return 1;
}

int senk_csr2csc(const double *val, const int *cind, const int *rptr, double **cval, int **crind, int **ccptr, int N,
                 int M) {
    if (!val || !cind || !rptr || !cval || !crind || !ccptr || N < 0 || M < 0) return SENK_ERR_ARG;
    const int nnz = rptr[N];
    std::vector<int> cp((size_t)M + 1, 0);
    for (int j = 0; j < nnz; j++) cp[cind[j] + 1]++;
    for (int c = 0; c < M; c++) cp[c + 1] += cp[c];
    std::vector<int> ri((size_t)nnz), fill(cp.begin(), cp.end() - 1);
    std::vector<double> cv((size_t)nnz);
    for (int i = 0; i < N; i++)
        for (int j = rptr[i]; j < rptr[i + 1]; j++) {
            const int o = fill[cind[j]]++;
            ri[o] = i;
            cv[o] = val[j];
        }
    *cval = dup_malloc(cv);
    *crind = dup_malloc(ri);
    *ccptr = dup_malloc(cp);
    return (*cval && *crind && *ccptr) ? SENK_OK : SENK_ERR_ALLOC;
}

// Csr2Sell (senk_matrix.hpp:90-142): slices of `size` rows, slice width = longest row,
// slot-major with the slice's own height as stride, padding 0 at column 0.
int senk_csr2sell(const double *val, const int *cind, const int *rptr, double **s_val, int **s_cind, int **s_len,
                  int size, int N, int *nnz) {
    if (!val || !cind || !rptr || !s_val || !s_cind || !s_len || size <= 0 || N < 0) return SENK_ERR_ARG;
    const int ns = (N + size - 1) / size;
    std::vector<int> wid((size_t)ns + 1, 0);
    long long total = 0;
    for (int s = 0; s < ns; s++) {
        const int h = (s == ns - 1 && N % size) ? N % size : size;
        int w = 0;
        for (int k = 0; k < h; k++) w = std::max(w, rptr[s * size + k + 1] - rptr[s * size + k]);
        wid[s + 1] = wid[s] + w;
        total += (long long)w * h;
    }
    std::vector<double> sv((size_t)total);
    std::vector<int> sc((size_t)total);
    for (int s = 0; s < ns; s++) {
        const int h = (s == ns - 1 && N % size) ? N % size : size;
        const int w = wid[s + 1] - wid[s];
        const size_t base = (size_t)wid[s] * size;
        for (int k = 0; k < h; k++) {
            const int r = s * size + k, n = rptr[r + 1] - rptr[r];
            for (int q = 0; q < w; q++) {
                const size_t o = base + (size_t)q * h + k;
                sv[o] = q < n ? val[rptr[r] + q] : 0.0;
                sc[o] = q < n ? cind[rptr[r] + q] : 0;
            }
        }
    }
    *s_val = dup_malloc(sv);
    *s_cind = dup_malloc(sc);
    *s_len = dup_malloc(wid);
    if (nnz) *nnz = (int)total;
    return SENK_OK;
}

// SELL-C-sigma packing: stable sort by decreasing row length inside windows of sigma
// rows, slices of C packed rows, slot-major with stride C; padding slots hold value 0
// at column 0 as in Csr2Sell, so sigma = 1 with N % C == 0 reproduces Csr2Sell's arrays.
int senk_csr2sellcs(const double *val, const int *cind, const int *rptr, int N, int C, int sigma, senk_sellcs *out) {
    if (!cind || !rptr || !out || C <= 0 || sigma <= 0 || N < 0) return SENK_ERR_ARG;
    std::memset(out, 0, sizeof(*out));
    const int ns = (N + C - 1) / C;
    std::vector<int> order((size_t)ns * C, -1);
    for (int i = 0; i < N; i++) order[i] = i;
    for (int w0 = 0; w0 < N; w0 += sigma) {
        const int w1 = std::min(N, w0 + sigma);
        std::stable_sort(order.begin() + w0, order.begin() + w1, [&](int a, int b) {
            return (rptr[a + 1] - rptr[a]) > (rptr[b + 1] - rptr[b]);
        });
    }
    std::vector<int> sp((size_t)ns + 1, 0);
    for (int s = 0; s < ns; s++) {
        int w = 0;
        for (int k = 0; k < C; k++) {
            const int r = order[(size_t)s * C + k];
            if (r >= 0) w = std::max(w, rptr[r + 1] - rptr[r]);
        }
        sp[s + 1] = sp[s] + w;
    }
    const size_t total = (size_t)sp[ns] * C;
    out->N = N;
    out->C = C;
    out->sigma = sigma;
    out->nslices = ns;
    out->nnz_padded = (int)total;
    out->slice_ptr = dup_malloc(sp);
    out->perm = dup_malloc(order);
    out->cind = (int *)std::malloc(sizeof(int) * (total ? total : 1));
    out->val64 = (double *)std::malloc(sizeof(double) * (total ? total : 1));
    out->val32 = nullptr;
    if (!out->slice_ptr || !out->perm || !out->cind || !out->val64) return SENK_ERR_ALLOC;
    for (int s = 0; s < ns; s++) {
        const int w = sp[s + 1] - sp[s];
        const size_t base = (size_t)sp[s] * C;
        for (int k = 0; k < C; k++) {
            const int r = order[(size_t)s * C + k];
            const int n = r >= 0 ? rptr[r + 1] - rptr[r] : 0;
            for (int q = 0; q < w; q++) {
                const size_t o = base + (size_t)q * C + k;
                out->cind[o] = q < n ? cind[rptr[r] + q] : 0;
                out->val64[o] = (q < n && val) ? val[rptr[r] + q] : 0.0;
            }
        }
    }
    return SENK_OK;
}

void senk_sellcs_free_host(senk_sellcs *s) {
    if (!s) return;
    std::free(s->slice_ptr);
    std::free(s->cind);
    std::free(s->perm);
    std::free(s->val64);
    std::free(s->val32);
    std::memset(s, 0, sizeof(*s));
}

// Csr2Bcsr (senk_matrix.hpp:144-211): blocks of a block row are the distinct column
// blocks col/bnw of its bnl rows, ascending; values column-major inside the block.
int senk_csr2bcsr(const double *val, const int *cind, const int *rptr, double **bval, int **bcind, int **brptr, int N,
                  int bnl, int bnw) {
    if (!val || !cind || !rptr || !bval || !bcind || !brptr || bnl <= 0 || bnw <= 0 || N < 0) return SENK_ERR_ARG;
    if (N % bnl || N % bnw) return SENK_ERR_DIVISIBILITY;
    const int ng = N / bnl, bs = bnl * bnw;
    std::vector<int> bp((size_t)ng + 1, 0), bc;
    std::vector<int> tmp;
    for (int g = 0; g < ng; g++) {
        tmp.clear();
        for (int k = 0; k < bnl; k++)
            for (int j = rptr[g * bnl + k]; j < rptr[g * bnl + k + 1]; j++) tmp.push_back(cind[j] / bnw);
        std::sort(tmp.begin(), tmp.end());
        tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
        bc.insert(bc.end(), tmp.begin(), tmp.end());
        bp[g + 1] = (int)bc.size();
    }
    std::vector<double> bv(bc.size() * (size_t)bs, 0.0);
    for (int g = 0; g < ng; g++)
        for (int k = 0; k < bnl; k++) {
            int b = bp[g];
            for (int j = rptr[g * bnl + k]; j < rptr[g * bnl + k + 1]; j++) {
                const int cb = cind[j] / bnw;
                while (bc[b] != cb) b++;
                bv[(size_t)b * bs + (size_t)(cind[j] % bnw) * bnl + k] = val[j];
            }
        }
    *bval = dup_malloc(bv);
    *bcind = dup_malloc(bc);
    *brptr = dup_malloc(bp);
    return SENK_OK;
}

// Bcsr2Csr (senk_matrix.hpp:213-241): every slot of every block, zeros included.
int senk_bcsr2csr(const double *bval, const int *bcind, const int *brptr, double **val, int **cind, int **rptr, int N,
                  int bnl, int bnw, int *nnz) {
    if (!bval || !bcind || !brptr || !val || !cind || !rptr || bnl <= 0 || bnw <= 0) return SENK_ERR_ARG;
    if (N % bnl) return SENK_ERR_DIVISIBILITY;
    const int bs = bnl * bnw;
    std::vector<int> rp((size_t)N + 1, 0), ci;
    std::vector<double> v;
    ci.reserve((size_t)brptr[N / bnl] * bs);
    v.reserve((size_t)brptr[N / bnl] * bs);
    for (int i = 0; i < N; i++) {
        const int g = i / bnl, k = i % bnl;
        for (int b = brptr[g]; b < brptr[g + 1]; b++)
            for (int l = 0; l < bnw; l++) {
                ci.push_back(bcind[b] * bnw + l);
                v.push_back(bval[(size_t)b * bs + (size_t)l * bnl + k]);
            }
        rp[i + 1] = (int)ci.size();
    }
    *val = dup_malloc(v);
    *cind = dup_malloc(ci);
    *rptr = dup_malloc(rp);
    if (nnz) *nnz = rp[N];
    return SENK_OK;
}

// Padding (senk_matrix.hpp:243-255): identity rows appended to a multiple of size;
// the arrays are reallocated.
int senk_padding(double **val, int **cind, int **rptr, int size, int *N) {
    if (!val || !cind || !rptr || !N || size <= 0) return SENK_ERR_ARG;
    const int n = *N, extra = n % size ? size - n % size : 0;
    if (!extra) return SENK_OK;
    const int nnz = (*rptr)[n];
    double *v = (double *)std::realloc(*val, sizeof(double) * (size_t)(nnz + extra));
    int *c = (int *)std::realloc(*cind, sizeof(int) * (size_t)(nnz + extra));
    int *r = (int *)std::realloc(*rptr, sizeof(int) * (size_t)(n + extra + 1));
    if (!v || !c || !r) return SENK_ERR_ALLOC;
    for (int q = 0; q < extra; q++) {
        v[nnz + q] = 1.0;
        c[nnz + q] = n + q;
        r[n + 1 + q] = r[n + q] + 1;
    }
    *val = v;
    *cind = c;
    *rptr = r;
    *N = n + extra;
    return SENK_OK;
}

// Split (senk_matrix.hpp:257-333)
int senk_split(const double *val, const int *cind, const int *rptr, double **lval, int **lcind, int **lrptr,
               double **uval, int **ucind, int **urptr, double **diag, int N, const char *key, int invDiag) {
    if (!val || !cind || !rptr || !key || !lval || !lcind || !lrptr) return SENK_ERR_ARG;
    int k;
    if (!std::strcmp(key, "LD-U")) k = 0;
    else if (!std::strcmp(key, "L-DU")) k = 1;
    else if (!std::strcmp(key, "L-D-U")) k = 2;
    else if (!std::strcmp(key, "LU-D")) k = 3;
    else return SENK_ERR_ARG;
    if (k != 3 && (!uval || !ucind || !urptr)) return SENK_ERR_ARG;
    if (k >= 2 && !diag) return SENK_ERR_ARG;
    std::vector<double> lv, uv, dv;
    std::vector<int> lc, uc, lr((size_t)N + 1, 0), ur((size_t)N + 1, 0);
    for (int i = 0; i < N; i++) {
        for (int j = rptr[i]; j < rptr[i + 1]; j++) {
            const int c = cind[j];
            if (c < i || (c > i && k == 3)) { lv.push_back(val[j]); lc.push_back(c); }
            else if (c > i) { uv.push_back(val[j]); uc.push_back(c); }
            else {
                const double d = invDiag ? 1 / val[j] : val[j];
                if (k == 0) { lv.push_back(d); lc.push_back(c); }
                else if (k == 1) { uv.push_back(d); uc.push_back(c); }
                else dv.push_back(d);
            }
        }
        lr[i + 1] = (int)lv.size();
        ur[i + 1] = (int)uv.size();
    }
    *lval = dup_malloc(lv);
    *lcind = dup_malloc(lc);
    *lrptr = dup_malloc(lr);
    if (k != 3) {
        *uval = dup_malloc(uv);
        *ucind = dup_malloc(uc);
        *urptr = dup_malloc(ur);
    }
    if (k >= 2) *diag = dup_malloc(dv);
    return SENK_OK;
}

int senk_get_diag(const double *val, const int *cind, const int *rptr, double **diag, int N) {
    if (!val || !cind || !rptr || !diag || N < 0) return SENK_ERR_ARG;
    double *d = (double *)std::malloc(sizeof(double) * (N ? N : 1));
    if (!d) return SENK_ERR_ALLOC;
    for (int i = 0; i < N; i++)
        for (int j = rptr[i]; j < rptr[i + 1]; j++)
            if (cind[j] == i) { d[i] = val[j]; break; }
    *diag = d;
    return SENK_OK;
}

int senk_scaling(double *val, const int *cind, const int *rptr, const double *ld, const double *rd, int N) {
    if (!val || !cind || !rptr || !ld || !rd) return SENK_ERR_ARG;
    for (int i = 0; i < N; i++)
        for (int j = rptr[i]; j < rptr[i + 1]; j++) val[j] = ld[i] * val[j] * rd[cind[j]];
    return SENK_OK;
}

// AllocLevelZero (senk_matrix.hpp:551-650).  Level-of-fill symbolic factorisation with a
// sorted linked list per row: pivots are taken in increasing column, a pivot k of
// level lk <= p offers every column j after the diagonal of row k at level
// lk + lev(k,j) + 1, an existing column keeps the smaller level, and a column whose
// level exceeds p is neither a pivot nor kept.  This is the reference's work-row
// rebuild without the copies (its level reset on used pivots only touches entries
// left of the diagonal, which are never read again).  Rows need their diagonal.
int senk_alloc_level_zero(double **val, int **cind, int **rptr, int N, int p) {
    if (!val || !cind || !rptr || !*val || !*cind || !*rptr || N < 0 || p < 0) return SENK_ERR_ARG;
    const int *rp = *rptr, *ci = *cind;
    const double *v = *val;
    if (!senk_check_structure(ci, rp, N)) return SENK_ERR_ARG;
    std::vector<int> nr((size_t)N + 1, 0), nc, nl, ndiag((size_t)N, -1);
    std::vector<double> nv;
    nc.reserve((size_t)rp[N] * 2);
    nl.reserve((size_t)rp[N] * 2);
    nv.reserve((size_t)rp[N] * 2);
    std::vector<int> next((size_t)N + 1, -1), lev((size_t)N, 0), pos((size_t)N, -1);
    std::vector<int> touched;
    const int HEAD = N;  // list head sentinel
    for (int i = 0; i < N; i++) {
        // list of the row's original columns
        touched.clear();
        int prev = HEAD;
        for (int j = rp[i]; j < rp[i + 1]; j++) {
            const int c = ci[j];
            next[prev] = c;
            prev = c;
            lev[c] = 0;
            pos[c] = j;
            touched.push_back(c);
        }
        next[prev] = -1;
        for (int k = next[HEAD]; k != -1 && k < i; k = next[k]) {
            if (lev[k] > p) continue;
            const int lk = lev[k];
            int cur = k;  // insertion cursor: the list is ascending
            for (int q = ndiag[k] + 1; q < nr[k + 1]; q++) {
                const int c = nc[q], cand = lk + nl[q] + 1;
                while (next[cur] != -1 && next[cur] < c) cur = next[cur];
                if (next[cur] == c) {
                    if (lev[c] > cand) lev[c] = cand;
                } else {
                    next[c] = next[cur];
                    next[cur] = c;
                    lev[c] = cand;
                    pos[c] = -1;
                    touched.push_back(c);
                }
                cur = c;
            }
        }
        for (int c = next[HEAD]; c != -1; c = next[c]) {
            if (lev[c] > p) continue;
            if (c == i) ndiag[i] = (int)nc.size();
            nc.push_back(c);
            nl.push_back(lev[c]);
            nv.push_back(pos[c] >= 0 ? v[pos[c]] : 0.0);
        }
        nr[i + 1] = (int)nc.size();
        for (int c : touched) { pos[c] = -1; lev[c] = 0; }
    }
    double *ov = dup_malloc(nv);
    int *oc = dup_malloc(nc), *orp = dup_malloc(nr);
    if (!ov || !oc || !orp) return SENK_ERR_ALLOC;
    std::free(*val);
    std::free(*cind);
    std::free(*rptr);
    *val = ov;
    *cind = oc;
    *rptr = orp;
    return SENK_OK;
}

// AllocBlockZero (senk_matrix.hpp:652-666): every row of a bnl-row group gets every
// column of every bnw-column block touched by the group; new slots are 0.
int senk_alloc_block_zero(double **val, int **cind, int **rptr, int N, int bnl, int bnw) {
    if (!val || !cind || !rptr || bnl <= 0 || bnw <= 0) return SENK_ERR_ARG;
    double *bv;
    int *bc, *bp;
    SENK_TRY(senk_csr2bcsr(*val, *cind, *rptr, &bv, &bc, &bp, N, bnl, bnw));
    double *v;
    int *c, *r, nnz;
    const int st = senk_bcsr2csr(bv, bc, bp, &v, &c, &r, N, bnl, bnw, &nnz);
    std::free(bv);
    std::free(bc);
    std::free(bp);
    if (st != SENK_OK) return st;
    std::free(*val);
    std::free(*cind);
    std::free(*rptr);
    *val = v;
    *cind = c;
    *rptr = r;
    return SENK_OK;
}

// RemoveOffDiagonal (senk_matrix.hpp:668-689)
int senk_remove_off_diagonal(double **val, int **cind, int **rptr, int N, int bnum) {
    if (!val || !cind || !rptr || bnum <= 0 || N < 0) return SENK_ERR_ARG;
    const int bs = N / bnum;
    if (bs <= 0) return SENK_ERR_ARG;
    int *rp = *rptr, *ci = *cind;
    double *v = *val;
    int n = 0, start = rp[0];
    for (int i = 0; i < N; i++) {
        const int lo = i / bs * bs, hi = lo + bs, end = rp[i + 1];
        for (int j = start; j < end; j++)
            if (ci[j] >= lo && ci[j] < hi) { v[n] = v[j]; ci[n] = ci[j]; n++; }
        start = end;
        rp[i + 1] = n;
    }
    return SENK_OK;
}

namespace {
void sort_row_pairs(int *c, double *v, int n, std::vector<std::pair<int, double>> &tmp) {
    if (n < 2) return;
    if (n <= 32) {
        for (int a = 1; a < n; a++) {
            const int kc = c[a];
            const double kv = v[a];
            int b = a - 1;
            while (b >= 0 && c[b] > kc) { c[b + 1] = c[b]; v[b + 1] = v[b]; b--; }
            c[b + 1] = kc;
            v[b + 1] = kv;
        }
        return;
    }
    tmp.resize((size_t)n);
    for (int a = 0; a < n; a++) tmp[a] = {c[a], v[a]};
    std::sort(tmp.begin(), tmp.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
    for (int a = 0; a < n; a++) { c[a] = tmp[a].first; v[a] = tmp[a].second; }
}
}  // namespace

// Reordering (senk_matrix.hpp:691-701): column relabel and per-row sort.
int senk_reordering_rp(double *val, int *cind, const int *rptr, const int *RP, int N) {
    if (!val || !cind || !rptr || !RP) return SENK_ERR_ARG;
    std::vector<std::pair<int, double>> tmp;
    for (int i = 0; i < N; i++) {
        for (int j = rptr[i]; j < rptr[i + 1]; j++) cind[j] = RP[cind[j]];
        sort_row_pairs(cind + rptr[i], val + rptr[i], rptr[i + 1] - rptr[i], tmp);
    }
    return SENK_OK;
}
// Reordering (senk_matrix.hpp:703-730): B = P A P^T.
int senk_reordering(double *val, int *cind, int *rptr, const int *LP, const int *RP, int N) {
    if (!val || !cind || !rptr || !LP || !RP) return SENK_ERR_ARG;
    const int nnz = rptr[N];
    std::vector<double> tv((size_t)nnz);
    std::vector<int> tc((size_t)nnz), tr((size_t)N + 1, 0);
    for (int i = 0; i < N; i++) tr[i + 1] = tr[i] + rptr[LP[i] + 1] - rptr[LP[i]];
    std::vector<std::pair<int, double>> tmp;
    for (int i = 0; i < N; i++) {
        const int o = LP[i], len = rptr[o + 1] - rptr[o];
        for (int q = 0; q < len; q++) {
            tv[tr[i] + q] = val[rptr[o] + q];
            tc[tr[i] + q] = RP[cind[rptr[o] + q]];
        }
        sort_row_pairs(tc.data() + tr[i], tv.data() + tr[i], len, tmp);
    }
    std::memcpy(val, tv.data(), sizeof(double) * (size_t)nnz);
    std::memcpy(cind, tc.data(), sizeof(int) * (size_t)nnz);
    std::memcpy(rptr, tr.data(), sizeof(int) * ((size_t)N + 1));
    return SENK_OK;
}

int senk_amc_permutation(const int *cind, const int *rptr, int *num_color, int **size_color, int **LP, int **RP,
                         int N, int isSym) {
    if (!cind || !rptr || !num_color || !size_color || !LP || !RP || N < 0) return SENK_ERR_ARG;
    std::vector<int> sc, lp, rp;
    const int nc = host::amc(cind, rptr, N, isSym != 0, sc, lp, rp);
    *num_color = nc;
    *size_color = dup_malloc(sc);
    *LP = dup_malloc(lp);
    *RP = dup_malloc(rp);
    return SENK_OK;
}

int senk_abmc_permutation(const int *cind, const int *rptr, int *num_color, int **size_color, int **LP, int **RP,
                          int N, int bsize, int isSym, const char *bmethod) {
    if (!cind || !rptr || !num_color || !size_color || !LP || !RP || !bmethod || bsize <= 0) return SENK_ERR_ARG;
    bool connect;
    if (!std::strcmp(bmethod, "simple")) connect = false;
    else if (!std::strcmp(bmethod, "connect")) connect = true;
    else return SENK_ERR_ARG;
    if (N % bsize) return SENK_ERR_DIVISIBILITY;
    std::vector<int> sc, lp, rp;
    const int nc = host::abmc(cind, rptr, N, bsize, isSym != 0, connect, sc, lp, rp);
    *num_color = nc;
    *size_color = dup_malloc(sc);
    *LP = dup_malloc(lp);
    *RP = dup_malloc(rp);
    return SENK_OK;
}

void senk_trsv(const double *U, const double *b, double *x, int n, int m) {
    for (int i = m - 1; i >= 0; i--) {
        double t = b[i];
        for (int j = m - 1; j > i; j--) t -= U[(size_t)j * n + i] * x[j];
        x[i] = t / U[(size_t)i * n + i];
    }
}
double senk_ggen(double a, double b, double *c, double *s) {
    const double r = std::sqrt(a * a + b * b);
    *c = a / r;
    *s = -b / r;
    return r;
}
void senk_grot(double c, double s, double *a, double *b) {
    const double t = *a;
    *a = c * t - s * *b;
    *b = s * t + c * *b;
}
void senk_host_free(void *p) { std::free(p); }

}  // extern "C"
