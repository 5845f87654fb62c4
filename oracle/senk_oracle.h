/* senk_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C CPU restatement of the reference SenK algorithms on the preconditioned
 * Krylov path.  It is the checker for the HIP library: only tests/, the smoke()
 * entry of __graft_entry__.py and bench.py's cpu_baseline leg may load it, and
 * never as the code being measured or shipped.
 *
 * Argument order follows the reference templates so that the parity tests read
 * like calls into the reference.  Each function cites the reference file:line it
 * restates.  Floating point is evaluated in source order without contraction
 * (built with -ffp-contract=off), exactly as the reference is compiled for the
 * x86-64 baseline, so sequential results are bitwise comparable with it.
 *
 * Pinning status (how each function is checked before it is trusted) is listed in
 * the header comment of senk_oracle.c.
 */
#ifndef SENK_ORACLE_H
#define SENK_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* ---- sparse kernels (senk_sparse.hpp) ---- */
void or_spmv_csr_d(const double *val, const int *cind, const int *rptr, const double *x, double *y, int N);
void or_spmv_csr_f(const float *val, const int *cind, const int *rptr, const float *x, float *y, int N);
void or_spmv_csr_diag_d(const double *val, const int *cind, const int *rptr, const double *diag,
                        const double *x, double *y, int N);
void or_spmv_bcsr_d(const double *bval, const int *bcind, const int *brptr, const double *x, double *y,
                    int N, int bnl, int bnw);
void or_spmv_bcsr_f(const float *bval, const int *bcind, const int *brptr, const float *x, float *y,
                    int N, int bnl, int bnw);
void or_spmv_sell_d(const double *val, const int *cind, const int *wid, int len, const double *x, double *y, int N);
void or_spmv_sell_f(const float *val, const int *cind, const int *wid, int len, const float *x, float *y, int N);

void or_sptrsv_l_d(const double *val, const int *cind, const int *rptr, const double *x, double *y, int N);
void or_sptrsv_u_d(const double *val, const int *cind, const int *rptr, const double *x, double *y, int N);
void or_sptrsv_l_f(const float *val, const int *cind, const int *rptr, const float *x, float *y, int N);
void or_sptrsv_u_f(const float *val, const int *cind, const int *rptr, const float *x, float *y, int N);
void or_sptrsv_l_amc_d(const double *val, const int *cind, const int *rptr, const double *x, double *y,
                       int N, const int *cptr, int cnum);
void or_sptrsv_u_amc_d(const double *val, const int *cind, const int *rptr, const double *x, double *y,
                       int N, const int *cptr, int cnum);
void or_sptrsv_l_abmc_d(const double *val, const int *cind, const int *rptr, const double *x, double *y,
                        int N, const int *cptr, int cnum, int bsize);
void or_sptrsv_u_abmc_d(const double *val, const int *cind, const int *rptr, const double *x, double *y,
                        int N, const int *cptr, int cnum, int bsize);
void or_sptrsv_l_bj_d(const double *val, const int *cind, const int *rptr, const double *x, double *y, int N, int bnum);
void or_sptrsv_u_bj_d(const double *val, const int *cind, const int *rptr, const double *x, double *y, int N, int bnum);
void or_sptrsv_bcsr_l_d(const double *bval, const int *bcind, const int *brptr, const double *x, double *y,
                        int N, int bnl, int bnw);
void or_sptrsv_bcsr_u_d(const double *bval, const int *bcind, const int *brptr, const double *x, double *y,
                        int N, int bnl, int bnw);
void or_sptrsv_bcsr_l_abmc_d(const double *bval, const int *bcind, const int *brptr, const double *x, double *y,
                             int N, const int *cptr, int cnum, int bsize, int bnl, int bnw);
void or_sptrsv_bcsr_u_abmc_d(const double *bval, const int *bcind, const int *brptr, const double *x, double *y,
                             int N, const int *cptr, int cnum, int bsize, int bnl, int bnw);

/* ---- BLAS-1 / BLAS-2 (senk_blas1.hpp, senk_blas2.hpp) ---- */
void or_copy_d(const double *x, double *y, int N);
void or_scal_d(double a, double *x, int N);
void or_axpy_d(double a, const double *x, double *y, int N);
void or_axpby_d(double a, const double *x, double b, double *y, int N);
void or_axpyz_d(double a, const double *x, const double *y, double *z, int N);
double or_dot_d(const double *x, const double *y, int N);
double or_nrm2_d(const double *x, int N);
void or_hadprod_d(const double *x, double *y, int N);
void or_haddiv_d(const double *x, double *y, int N);
double or_ggen_d(double a, double b, double *c, double *s);
void or_grot_d(double c, double s, double *a, double *b);
void or_trsv_d(const double *U, const double *b, double *x, int n, int m);

/* ---- matrix setup (senk_matrix.hpp) ---- */
int or_check_structure(const int *cind, const int *rptr, int N);
void or_csr2csc_d(const double *val, const int *cind, const int *rptr, double *cval, int *crind, int *ccptr,
                  int N, int M);
int or_csr2sell_count(const int *rptr, int *s_len, int size, int N);
void or_csr2sell_fill_d(const double *val, const int *cind, const int *rptr, const int *s_len,
                        double *s_val, int *s_cind, int size, int N);
int or_csr2bcsr_count(const int *cind, const int *rptr, int *brptr, int N, int bnl, int bnw);
void or_csr2bcsr_fill_d(const double *val, const int *cind, const int *rptr, const int *brptr,
                        double *bval, int *bcind, int N, int bnl, int bnw);
int or_bcsr2csr_d(const double *bval, const int *bcind, const int *brptr, double *val, int *cind, int *rptr,
                  int N, int bnl, int bnw);
int or_padding_d(double *val, int *cind, int *rptr, int size, int N);
int or_split_d(const double *val, const int *cind, const int *rptr,
               double *lval, int *lcind, int *lrptr, double *uval, int *ucind, int *urptr,
               double *diag, int N, int key, int invDiag);
void or_get_diag_d(const double *val, const int *cind, const int *rptr, double *diag, int N);
void or_scaling_d(double *val, const int *cind, const int *rptr, const double *ld, const double *rd, int N);
int or_ilu0_d(double *val, const int *cind, const int *rptr, int N);
int or_ilu0_f(float *val, const int *cind, const int *rptr, int N);
int or_ic0_d(double *val, const int *cind, const int *rptr, int N);
int or_alloc_level_zero_d(const double *val, const int *cind, const int *rptr, double *oval, int *ocind, int *orptr,
                          int N, int p, int cap);
int or_alloc_block_zero_d(const double *val, const int *cind, const int *rptr, double *oval, int *ocind, int *orptr,
                          int N, int bnl, int bnw, int cap);
int or_ilup_d(const double *val, const int *cind, const int *rptr, double *oval, int *ocind, int *orptr,
              int N, int p, int cap);
int or_remove_off_diagonal_d(double *val, int *cind, int *rptr, int N, int bnum);
void or_reordering_rp_d(double *val, int *cind, const int *rptr, const int *RP, int N);
void or_reordering_d(double *val, int *cind, int *rptr, const int *LP, const int *RP, int N);

/* ---- graph ordering (senk_graph.hpp) ---- */
int or_coloring(const int *cind, const int *rptr, int *color, int N);
int or_adjacency(const int *cind, const int *rptr, int *am_cind, int *am_rptr, int N, int isSym);
int or_amc_perm(const int *cind, const int *rptr, int *size_color, int *LP, int *RP, int N, int isSym);
int or_abmc_perm(const int *cind, const int *rptr, int *size_color, int *LP, int *RP, int N, int bsize,
                 int isSym, int method);

/* ---- Krylov drivers (senk_gmres.hpp, senk_gcr.hpp, senk_bicgstab.hpp) ----
 * prec: 0 none; 1 ILU from CSR factors (L strict lower, U with inverted diagonal first);
 *       2 ILUB from BCSR factors (bnl x bnw).  BCSR factors go in the lval/uval slots.
 * hist receives the residual the reference prints/tests each step (|e_{j+1}|/nrm_b for
 * GMRES, ||r||/nrm_b for BiCGStab and at GCR restarts).  Returns the iteration count the
 * reference reports; *converged is set to 1 when the stop test fired. */
typedef struct {
    int prec, bnl, bnw;
    const double *lval; const int *lcind; const int *lrptr;
    const double *uval; const int *ucind; const int *urptr;
} or_prec_t;
int or_gmres_d(const double *val, const int *cind, const int *rptr, const or_prec_t *M,
               const double *b, double *x, double nrm_b, int outer, int m, int N, double eps,
               double *hist, int hist_cap, int *converged);
int or_gcr_d(const double *val, const int *cind, const int *rptr, const or_prec_t *M,
             const double *b, double *x, double nrm_b, int outer, int m, int N, double eps,
             double *hist, int hist_cap, int *converged);
int or_bicgstab_d(const double *val, const int *cind, const int *rptr, const or_prec_t *M,
                  const double *b, double *x, double nrm_b, int max_iter, int N, double eps,
                  double *hist, int hist_cap, int *converged);
int or_cg_d(const double *val, const int *cind, const int *rptr, const or_prec_t *M, const double *b, double *x,
            double nrm_b, int max_iter, int N, double eps, double *hist, int hist_cap, int *converged);
double or_relative_residual_d(const double *val, const int *cind, const int *rptr, const double *b,
                              const double *x, int N);

#ifdef __cplusplus
}
#endif
#endif
