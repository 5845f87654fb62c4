// Driver for tests/test_host_sanitizers.py: reads every file named on the command line
// with senk_read_matrix_market / senk_matrix_market_size, expands symmetric ones, drops
// zeros, frees everything, and prints one status line per file.
#include <cstdio>
#include <cstdlib>
#include <string>

#include "senk_amd.h"

namespace senk_amd {
void record_error(const std::string &) {}
}  // namespace senk_amd

int main(int argc, char **argv) {
    for (int a = 1; a < argc; a++) {
        double *v;
        int *c, *r, N, M, sh, n2, m2, sh2;
        long long E;
        const int st = senk_read_matrix_market(argv[a], &v, &c, &r, &N, &M, &sh, 1);
        const int st2 = senk_matrix_market_size(argv[a], &n2, &m2, &E, &sh2);
        int nnz = -1, se = -1, sz = -1;
        if (st == SENK_OK) {
            nnz = r[N];
            if (sh == SENK_SHAPE_SYM && N == M) {
                double *ev;
                int *ec, *er;
                se = senk_expand(v, c, r, &ev, &ec, &er, N, "L");
                if (se == SENK_OK) {
                    std::free(ev);
                    std::free(ec);
                    std::free(er);
                }
            }
            sz = senk_remove_zeros(&v, &c, r, N);
            std::free(v);
            std::free(c);
            std::free(r);
        }
        std::printf("%d %d %d %d %d\n", st, st2, nnz, se, sz);
    }
    return 0;
}
