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
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include "internal.hpp"

namespace {

template <typename T>
T *malloc_n(size_t n) {
    return (T *)std::malloc(sizeof(T) * (n ? n : 1));
}

// The C locale's white space (what stream extraction skips): ' ', \t \n \v \f \r.
inline bool is_space(char ch) { return ch == ' ' || (ch >= '\t' && ch <= '\r'); }

// Whitespace-separated token reader over a whole file held in memory.
struct Cursor {
    const char *p, *end;
    void skip_space() {
        while (p < end && is_space(*p)) p++;
    }
    bool read_long(long long *v) {
        skip_space();
        if (p < end && *p == '+') p++;
        auto r = std::from_chars(p, end, *v);
        if (r.ec != std::errc()) return false;
        p = r.ptr;
        return p == end || is_space(*p);
    }
    bool read_double(double *v) {
        skip_space();
        if (p < end && *p == '+') p++;
        auto r = std::from_chars(p, end, *v, std::chars_format::general);
        if (r.ec != std::errc()) return false;
        p = r.ptr;
        return p == end || is_space(*p);
    }
};

// A whole file in memory: mapped read-only, or read into a buffer where mapping fails.
struct FileText {
    const char *data = nullptr;
    size_t size = 0;
    void *map = nullptr;
    std::unique_ptr<char[]> buf;
    ~FileText() {
        if (map) munmap(map, size);
    }
    // 0, or errno-style failure
    int open(const char *path) {
        const int fd = ::open(path, O_RDONLY);
        if (fd < 0) return 1;
        struct stat st;
        if (fstat(fd, &st) != 0) {
            ::close(fd);
            return 1;
        }
        size = (size_t)st.st_size;
        if (size > 0) {
            void *m = mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
            if (m != MAP_FAILED) {
                map = m;
                data = (const char *)m;
            } else {
                buf.reset(new char[size]);
                size_t got = 0;
                while (got < size) {
                    const ssize_t k = ::read(fd, buf.get() + got, size - got);
                    if (k <= 0) break;
                    got += (size_t)k;
                }
                size = got;
                data = buf.get();
            }
        }
        ::close(fd);
        return 0;
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

// The entry list: `pe` triples "row col value" as whitespace-separated tokens (a triple
// may span lines, as with stream extraction).  The text is cut into chunks at whitespace,
// each chunk's tokens are counted in parallel, a prefix sum gives every chunk the global
// index of its first token (and so the row / column / value phase of each token), and the
// chunks are parsed in parallel.  Returns the index of the first bad entry in file order
// (unparsable or outside n x m), or pe when all are good; *bad_range tells which.
long long parse_entries(const char *p, const char *end, long long pe, long long n, long long m, int *row, int *col,
                        double *val, bool *bad_range) {
    const long long need = 3 * pe;
    const size_t bytes = (size_t)(end - p);
    int nt = (int)std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency()));
    if (bytes < (size_t)(1 << 20)) nt = 1;
    std::vector<const char *> cut((size_t)nt + 1);
    cut[0] = p;
    cut[nt] = end;
    for (int t = 1; t < nt; t++) {
        const char *b = std::max(cut[t - 1], p + bytes / nt * t);
        while (b < end && !is_space(*b)) b++;  // never inside a token
        cut[t] = b;
    }
    std::vector<long long> ntok((size_t)nt + 1, 0), first((size_t)nt, need), range((size_t)nt, 0);
    auto run = [&](auto &&fn) {
        std::vector<std::thread> th;
        for (int t = 1; t < nt; t++) th.emplace_back(fn, t);
        fn(0);
        for (auto &x : th) x.join();
    };
    run([&](int t) {
        long long k = 0;
        bool in = false;
        for (const char *q = cut[t]; q < cut[t + 1]; q++) {
            const bool sp = is_space(*q);
            if (!sp && !in) k++;
            in = !sp;
        }
        ntok[(size_t)t + 1] = k;
    });
    for (int t = 0; t < nt; t++) ntok[(size_t)t + 1] += ntok[t];
    run([&](int t) {
        long long g = ntok[t];
        Cursor c{cut[t], cut[t + 1]};
        while (g < need) {
            c.skip_space();
            if (c.p >= c.end) break;
            const long long e = g / 3;
            bool ok;
            if (g % 3 == 2) {
                ok = c.read_double(&val[e]);
            } else {
                long long v;
                ok = c.read_long(&v);
                if (ok && (v < 1 || v > (g % 3 ? m : n))) {
                    ok = false;
                    range[t] = 1;
                }
                if (ok) (g % 3 ? col : row)[e] = (int)(v - 1);
            }
            if (!ok) {
                first[t] = g;
                return;
            }
            g++;
        }
    });
    long long bad = need;
    for (int t = 0; t < nt; t++)
        if (first[t] < bad) {
            bad = first[t];
            *bad_range = range[t] != 0;
        }
    if (bad == need && ntok[nt] < need) {  // the list ends early
        bad = ntok[nt];
        *bad_range = false;
    }
    return bad / 3;
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
            while (b.p < b.end && !is_space(*b.p)) b.p++;
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
        while (q < e && is_space(*q)) q++;
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
            while (*q && is_space(*q)) q++;
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
    FileText text;
    if (text.open(filename)) return fail(SENK_ERR_IO, std::string("could not open ") + filename);
    Cursor c{text.data, text.data + text.size};
    int sh;
    long long n, m, pe;
    if (int st = read_head(c, &sh, &n, &m, &pe)) return st;
    std::vector<int> row((size_t)pe), col((size_t)pe);
    std::vector<double> v((size_t)pe);
    bool bad_range = false;
    const long long bad = parse_entries(c.p, c.end, pe, n, m, row.data(), col.data(), v.data(), &bad_range);
    if (bad < pe)
        return fail(SENK_ERR_FORMAT, "MatrixMarket: entry " + std::to_string(bad + 1) +
                                         (bad_range ? " lies outside the matrix" : " is not 'row col value'"));
    if (removeZeros) {  // drop stored zeros, keeping file order
        size_t k = 0;
        for (size_t e = 0; e < (size_t)pe; e++)
            if (v[e] != 0) {
                row[k] = row[e];
                col[k] = col[e];
                v[k] = v[e];
                k++;
            }
        row.resize(k);
        col.resize(k);
        v.resize(k);
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
