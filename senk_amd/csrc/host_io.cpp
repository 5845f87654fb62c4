// MatrixMarket ingestion and the CSR pre-processing that follows it in the reference's
// workflow (main.cpp:20-40): reading a coordinate file into CSR, dropping stored zeros,
// expanding a stored triangle of a symmetric matrix, copying a CSR matrix.  Host code;
// integer outputs are identical to the reference's for every file the reference reads
// correctly, and the values are the file's values (parsed correctly rounded).
#include <algorithm>
#include <cctype>
#include <charconv>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "internal.hpp"

namespace {

template <typename T>
T *malloc_n(size_t n) {
    return (T *)std::malloc(sizeof(T) * (n ? n : 1));
}

// Whitespace-separated token reader over a whole file held in memory.
struct Cursor {
    const char *p, *end;
    void skip_space() {
        while (p < end && std::isspace((unsigned char)*p)) p++;
    }
    bool read_long(long long *v) {
        skip_space();
        if (p < end && *p == '+') p++;
        auto r = std::from_chars(p, end, *v);
        if (r.ec != std::errc()) return false;
        p = r.ptr;
        return p == end || std::isspace((unsigned char)*p);
    }
    bool read_double(double *v) {
        skip_space();
        if (p < end && *p == '+') p++;
        auto r = std::from_chars(p, end, *v, std::chars_format::general);
        if (r.ec != std::errc()) return false;
        p = r.ptr;
        return p == end || std::isspace((unsigned char)*p);
    }
};

std::string lower(std::string s) {
    for (char &ch : s) ch = (char)std::tolower((unsigned char)ch);
    return s;
}

int fail(int code, const std::string &msg) {
    senk_amd::record_error(msg);
    return code;
}

// Banner ("%%MatrixMarket matrix coordinate <real|integer|double> <general|symmetric>",
// tokens case-insensitive), comment and blank lines, then "rows cols entries"; the
// cursor is left after the size line.  *sh: 0 symmetric, 1 general.
int read_head(Cursor &c, int *sh, long long *n, long long *m, long long *pe) {
    const char *eol = std::find(c.p, c.end, '\n');
    std::vector<std::string> tok;
    {
        Cursor b{c.p, eol};
        while (true) {
            b.skip_space();
            if (b.p >= b.end) break;
            const char *s = b.p;
            while (b.p < b.end && !std::isspace((unsigned char)*b.p)) b.p++;
            tok.push_back(lower(std::string(s, b.p)));
        }
    }
    if (tok.size() < 5 || tok[0] != "%%matrixmarket" || tok[1] != "matrix" || tok[2] != "coordinate")
        return fail(SENK_ERR_FORMAT, "MatrixMarket: banner is not '%%MatrixMarket matrix coordinate ...'");
    if (tok[3] != "real" && tok[3] != "integer" && tok[3] != "double")
        return fail(SENK_ERR_FORMAT, "MatrixMarket: field '" + tok[3] + "' is not real or integer");
    if (tok[4] == "general") *sh = 1;
    else if (tok[4] == "symmetric") *sh = 0;
    else return fail(SENK_ERR_FORMAT, "MatrixMarket: symmetry '" + tok[4] + "' is not general or symmetric");
    const char *p = eol < c.end ? eol + 1 : c.end;
    while (p < c.end) {
        const char *e = std::find(p, c.end, '\n');
        const char *q = p;
        while (q < e && std::isspace((unsigned char)*q)) q++;
        if (q < e && *q != '%') break;
        p = e < c.end ? e + 1 : c.end;
    }
    c.p = p;
    if (!c.read_long(n) || !c.read_long(m) || !c.read_long(pe) || *n < 0 || *m < 0 || *pe < 0)
        return fail(SENK_ERR_FORMAT, "MatrixMarket: size line is not 'rows cols entries'");
    if (*n >= INT_MAX || *m >= INT_MAX || *pe > INT_MAX)
        return fail(SENK_ERR_UNSUPPORTED, "MatrixMarket: sizes beyond 32-bit indices");
    return SENK_OK;
}

}  // namespace

extern "C" {

// The banner and size line of a MatrixMarket file, as the reference's reader reports
// them before it reads the entries (senk_io.hpp:52-68): rows, columns, stored entries
// and the shape (0 symmetric, 1 general).  Reads only the head of the file.
int senk_matrix_market_size(const char *filename, int *N, int *M, long long *entries, int *shape) {
    if (!filename || !N || !M || !entries || !shape) return SENK_ERR_ARG;
    FILE *f = std::fopen(filename, "rb");
    if (!f) return fail(SENK_ERR_IO, std::string("could not open ") + filename);
    // banner, comments and the size line: read lines until the first one that is
    // neither a comment nor blank after the banner
    std::string head;
    char line[4096];
    bool new_line = true, size_seen = false;
    long long lines = 0;
    while (std::fgets(line, sizeof line, f)) {
        const bool starts = new_line;  // a chunk that begins a line (long lines come in pieces)
        head += line;
        new_line = head.back() == '\n';
        if (starts && ++lines > 1) {
            const char *q = line;
            while (*q && std::isspace((unsigned char)*q)) q++;
            if (*q && *q != '%') size_seen = true;
        }
        if (size_seen && new_line) break;
    }
    std::fclose(f);
    Cursor c{head.data(), head.data() + head.size()};
    int sh;
    long long n, m, pe;
    if (int st = read_head(c, &sh, &n, &m, &pe)) return st;
    *N = (int)n;
    *M = (int)m;
    *entries = pe;
    *shape = sh;
    return SENK_OK;
}

// io::ReadMatrixMarket (senk_io.hpp:41-103).  The banner must name a real or integer
// coordinate matrix, "general" (*shape = 1, the reference's Unsym) or "symmetric"
// (*shape = 0, Sym: the file's triangle is returned as stored; senk_expand completes
// it).  removeZeros drops every stored entry whose value is zero (the reference's test
// is `t_val == 0`, diagonal included).  The output is CSR with rows in order and each
// row's columns ascending, malloc'd (senk_host_free).  Entries with equal coordinates,
// which the format does not allow, are kept in file order.
int senk_read_matrix_market(const char *filename, double **val, int **cind, int **rptr, int *N, int *M, int *shape,
                            int removeZeros) {
    if (!filename || !val || !cind || !rptr || !N || !M || !shape) return SENK_ERR_ARG;
    FILE *f = std::fopen(filename, "rb");
    if (!f) return fail(SENK_ERR_IO, std::string("could not open ") + filename);
    std::vector<char> buf;
    {
        char chunk[1 << 16];
        size_t got;
        while ((got = std::fread(chunk, 1, sizeof chunk, f)) > 0) buf.insert(buf.end(), chunk, chunk + got);
        const bool err = std::ferror(f);
        std::fclose(f);
        if (err) return fail(SENK_ERR_IO, std::string("could not read ") + filename);
    }
    Cursor c{buf.data(), buf.data() + buf.size()};
    int sh;
    long long n, m, pe;
    if (int st = read_head(c, &sh, &n, &m, &pe)) return st;
    std::vector<int> row, col;
    std::vector<double> v;
    row.reserve((size_t)pe);
    col.reserve((size_t)pe);
    v.reserve((size_t)pe);
    for (long long k = 0; k < pe; k++) {
        long long r, q;
        double x;
        if (!c.read_long(&r) || !c.read_long(&q) || !c.read_double(&x))
            return fail(SENK_ERR_FORMAT, "MatrixMarket: entry " + std::to_string(k + 1) + " is not 'row col value'");
        if (r < 1 || r > n || q < 1 || q > m)
            return fail(SENK_ERR_FORMAT, "MatrixMarket: entry " + std::to_string(k + 1) + " lies outside the matrix");
        if (removeZeros && x == 0) continue;
        row.push_back((int)(r - 1));
        col.push_back((int)(q - 1));
        v.push_back(x);
    }
    const size_t nnz = v.size();
    // two stable counting passes (by column, then by row): rows in order, columns
    // ascending in each row, equal coordinates in file order
    std::vector<int> cc((size_t)m + 1, 0), rp((size_t)n + 1, 0), ord(nnz), tmp(nnz);
    for (size_t k = 0; k < nnz; k++) cc[col[k] + 1]++;
    for (long long q = 0; q < m; q++) cc[q + 1] += cc[q];
    for (size_t k = 0; k < nnz; k++) tmp[cc[col[k]]++] = (int)k;
    for (size_t k = 0; k < nnz; k++) rp[row[k] + 1]++;
    for (long long r = 0; r < n; r++) rp[r + 1] += rp[r];
    {
        std::vector<int> fill(rp.begin(), rp.end() - 1);
        for (size_t k = 0; k < nnz; k++) ord[fill[row[tmp[k]]]++] = tmp[k];
    }
    double *ov = malloc_n<double>(nnz);
    int *oc = malloc_n<int>(nnz), *orp = malloc_n<int>((size_t)n + 1);
    if (!ov || !oc || !orp) {
        std::free(ov);
        std::free(oc);
        std::free(orp);
        return SENK_ERR_ALLOC;
    }
    for (size_t k = 0; k < nnz; k++) {
        ov[k] = v[ord[k]];
        oc[k] = col[ord[k]];
    }
    std::memcpy(orp, rp.data(), sizeof(int) * ((size_t)n + 1));
    *val = ov;
    *cind = oc;
    *rptr = orp;
    *N = (int)n;
    *M = (int)m;
    *shape = sh;
    return SENK_OK;
}

// RemoveZeros (senk_matrix.hpp:29-43): every entry whose value is zero leaves, in
// place; rptr is updated and val/cind are reallocated to the new length.
int senk_remove_zeros(double **val, int **cind, int *rptr, int N) {
    if (!val || !cind || !*val || !*cind || !rptr || N < 0) return SENK_ERR_ARG;
    double *v = *val;
    int *c = *cind;
    int out = rptr[0], start = rptr[0];
    for (int i = 0; i < N; i++) {
        const int stop = rptr[i + 1];
        for (int j = start; j < stop; j++)
            if (v[j] != 0) {
                v[out] = v[j];
                c[out] = c[j];
                out++;
            }
        start = stop;
        rptr[i + 1] = out;
    }
    double *nv = (double *)std::realloc(v, sizeof(double) * (out ? out : 1));
    int *nc = (int *)std::realloc(c, sizeof(int) * (out ? out : 1));
    if (nv) *val = nv;
    if (nc) *cind = nc;
    return (nv && nc) ? SENK_OK : SENK_ERR_ALLOC;
}

// Expand (senk_matrix.hpp:335-371): the full matrix from one stored triangle.  key "L"
// (the triangle holds the lower part with the diagonal, as a symmetric MatrixMarket
// file does): row i is the stored row i followed by column i of the stored matrix
// without its first entry (the diagonal).  key "U": column i of the stored matrix, then
// row i without its first entry.  Outputs are malloc'd; rptr[N] is the entry count
// (2 nnz - N when every diagonal entry is stored).
int senk_expand(const double *tval, const int *tcind, const int *trptr, double **val, int **cind, int **rptr, int N,
                const char *key) {
    if (!tval || !tcind || !trptr || !val || !cind || !rptr || !key || N < 0) return SENK_ERR_ARG;
    int lower_stored;
    if (!std::strcmp(key, "L")) lower_stored = 1;
    else if (!std::strcmp(key, "U")) lower_stored = 0;
    else return fail(SENK_ERR_ARG, "Expand: key is not \"L\" or \"U\"");
    const int nnz = trptr[N];
    for (int j = 0; j < nnz; j++)
        if (tcind[j] < 0 || tcind[j] >= N) return fail(SENK_ERR_ARG, "Expand: column index outside the matrix");
    // transpose (columns of the stored matrix), rows ascending in each column
    std::vector<int> tp((size_t)N + 1, 0), tr((size_t)nnz);
    std::vector<double> tv((size_t)nnz);
    for (int j = 0; j < nnz; j++) tp[tcind[j] + 1]++;
    for (int i = 0; i < N; i++) tp[i + 1] += tp[i];
    {
        std::vector<int> fill(tp.begin(), tp.end() - 1);
        for (int i = 0; i < N; i++)
            for (int j = trptr[i]; j < trptr[i + 1]; j++) {
                const int o = fill[tcind[j]]++;
                tr[o] = i;
                tv[o] = tval[j];
            }
    }
    // first part taken whole, second part without its first entry
    const int *ap = lower_stored ? trptr : tp.data();
    const int *ai = lower_stored ? tcind : tr.data();
    const double *av = lower_stored ? tval : tv.data();
    const int *bp = lower_stored ? tp.data() : trptr;
    const int *bi = lower_stored ? tr.data() : tcind;
    const double *bv = lower_stored ? tv.data() : tval;
    long long total = 0;
    for (int i = 0; i < N; i++) total += (ap[i + 1] - ap[i]) + std::max(0, bp[i + 1] - bp[i] - 1);
    if (total > INT_MAX) return fail(SENK_ERR_UNSUPPORTED, "Expand: more than 2^31-1 entries");
    double *ov = malloc_n<double>((size_t)total);
    int *oc = malloc_n<int>((size_t)total), *orp = malloc_n<int>((size_t)N + 1);
    if (!ov || !oc || !orp) {
        std::free(ov);
        std::free(oc);
        std::free(orp);
        return SENK_ERR_ALLOC;
    }
    int k = 0;
    orp[0] = 0;
    for (int i = 0; i < N; i++) {
        for (int j = ap[i]; j < ap[i + 1]; j++, k++) {
            ov[k] = av[j];
            oc[k] = ai[j];
        }
        for (int j = bp[i] + 1; j < bp[i + 1]; j++, k++) {
            ov[k] = bv[j];
            oc[k] = bi[j];
        }
        orp[i + 1] = k;
    }
    *val = ov;
    *cind = oc;
    *rptr = orp;
    return SENK_OK;
}

// Duplicate (senk_matrix.hpp:373-389): a malloc'd copy of a CSR matrix.
int senk_duplicate(const double *tval, const int *tcind, const int *trptr, double **val, int **cind, int **rptr,
                   int N) {
    if (!tval || !tcind || !trptr || !val || !cind || !rptr || N < 0) return SENK_ERR_ARG;
    const size_t nnz = (size_t)trptr[N];
    double *ov = malloc_n<double>(nnz);
    int *oc = malloc_n<int>(nnz), *orp = malloc_n<int>((size_t)N + 1);
    if (!ov || !oc || !orp) {
        std::free(ov);
        std::free(oc);
        std::free(orp);
        return SENK_ERR_ALLOC;
    }
    if (nnz) {
        std::memcpy(ov, tval, sizeof(double) * nnz);
        std::memcpy(oc, tcind, sizeof(int) * nnz);
    }
    std::memcpy(orp, trptr, sizeof(int) * ((size_t)N + 1));
    *val = ov;
    *cind = oc;
    *rptr = orp;
    return SENK_OK;
}

}  // extern "C"
