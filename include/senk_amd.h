/* senk_amd.h -- C ABI of the MI355X (gfx950) implementation of SenK's preconditioned
 * Krylov path.
 *
 * Conventions (mirroring the reference templates in include/senk_*.hpp):
 *  - argument order follows the reference function each entry point replaces; a
 *    leading senk_ctx selects the device and stream;
 *  - "d" pointers are device memory, everything else is host memory;
 *  - indices are int32 with the reference's meaning (CSR rptr/cind, BCSR brptr/bcind,
 *    ABMC cptr counted in blocks, AMC cptr counted in rows);
 *  - every entry point returns a status code instead of calling exit(): SENK_OK on
 *    success, and the code for the condition the reference reports with printf+exit;
 *  - host-side outputs whose size the reference decides (its T** arguments) are
 *    allocated with malloc() exactly as the reference does; release them with free()
 *    or senk_host_free();
 *  - work on one senk_ctx is stream-ordered; calls that return a host scalar or a
 *    result struct synchronise the context's stream.  A context must not be used from
 *    two host threads at once.
 */
#ifndef SENK_AMD_H
#define SENK_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
enum senk_status {
    SENK_OK = 0,
    SENK_ERR_ARG = 1,            /* invalid argument (null pointer, bad size, bad key)        */
    SENK_ERR_HIP = 2,            /* HIP runtime error                                          */
    SENK_ERR_ZERO_PIVOT = 3,     /* Ilu0 zero pivot (senk_matrix.hpp:426-429)                  */
    SENK_ERR_DIVISIBILITY = 4,   /* N % bsize, N % bnl or N % bnw (senk_graph.hpp:296,
                                    senk_matrix.hpp:150)                                      */
    SENK_ERR_BREAKDOWN = 5,      /* Krylov breakdown (non-finite scalar); the reference
                                    lets NaN propagate                                        */
    SENK_ERR_NOT_CONVERGED = 6,  /* reserved: solvers report non-convergence in the result   */
    SENK_ERR_NO_DEVICE = 7,      /* no HIP device                                              */
    SENK_ERR_ALLOC = 8,          /* allocation failure (senk_utils.hpp:26,38,51)               */
    SENK_ERR_UNSUPPORTED = 9,    /* shape outside what the kernels support                     */
    SENK_ERR_SCHEDULE = 10,      /* a factor entry couples two blocks scheduled together       */
    SENK_ERR_TIMEOUT = 11,       /* a single-launch triangular solve gave up waiting on an
                                    earlier level (the result of that solve is invalid)      */
    SENK_ERR_IO = 12,            /* a file could not be opened or read                         */
    SENK_ERR_FORMAT = 13         /* a file is not a MatrixMarket matrix the reader supports    */
};
const char *senk_status_string(int status);
/* Last HIP error text recorded by the library (thread-local), for diagnostics. */
const char *senk_last_error(void);
int senk_version(void);
void senk_host_free(void *p);

/* ------------------------------------------------------------------ context */
typedef struct senk_ctx_s *senk_ctx;
int senk_ctx_create(int device, senk_ctx *out);
int senk_ctx_destroy(senk_ctx ctx);
int senk_ctx_synchronize(senk_ctx ctx);
int senk_device_count(int *count);
int senk_device_malloc(senk_ctx ctx, size_t bytes, void **dptr);
int senk_device_free(senk_ctx ctx, void *dptr);
int senk_memcpy_htod(senk_ctx ctx, void *dst, const void *src, size_t bytes);
int senk_memcpy_dtoh(senk_ctx ctx, void *dst, const void *src, size_t bytes);
int senk_memcpy_dtod(senk_ctx ctx, void *dst, const void *src, size_t bytes);
int senk_memset_zero(senk_ctx ctx, void *dst, size_t bytes);
/* Stream-ordered timing: record a marker, later read the elapsed ms between two. */
int senk_event_create(senk_ctx ctx, void **ev);
int senk_event_destroy(void *ev);
int senk_event_record(senk_ctx ctx, void *ev);
int senk_event_elapsed_ms(void *start, void *stop, float *ms);

/* ------------------------------------------------------------------ sparse kernels
 * senk_sparse.hpp.  All arrays are device pointers. */

/* SpmvCsr<T>(val,cind,rptr,x,y,N)          senk_sparse.hpp:25-35 */
int senk_spmv_csr_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr,
                      const double *dx, double *dy, int N);
int senk_spmv_csr_f32(senk_ctx ctx, const float *dval, const int *dcind, const int *drptr,
                      const float *dx, float *dy, int N);
/* SpmvCsr<T>(val,cind,rptr,diag,x,y,N)     senk_sparse.hpp:47-57 */
int senk_spmv_csr_diag_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr,
                           const double *ddiag, const double *dx, double *dy, int N);
/* SpmvBcsr<T,bnl,bnw>(bval,bcind,brptr,x,y,N)   senk_sparse.hpp:70-91 */
int senk_spmv_bcsr_f64(senk_ctx ctx, const double *dbval, const int *dbcind, const int *dbrptr,
                       const double *dx, double *dy, int N, int bnl, int bnw);
int senk_spmv_bcsr_f32(senk_ctx ctx, const float *dbval, const int *dbcind, const int *dbrptr,
                       const float *dx, float *dy, int N, int bnl, int bnw);
/* SpmvSell<T>(val,cind,wid,len,x,y,N)      senk_sparse.hpp:103-121, on the layout written
 * by Csr2Sell (senk_matrix.hpp:90-142).  dwid is a device copy of the slice prefix. */
int senk_spmv_sell_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *dwid, int len,
                       const double *dx, double *dy, int N);
int senk_spmv_sell_f32(senk_ctx ctx, const float *dval, const int *dcind, const int *dwid, int len,
                       const float *dx, float *dy, int N);

/* SELL-C-sigma (no reference counterpart; extends Csr2Sell/SpmvSell with a sigma-window
 * row sort).  Rows are sorted by decreasing length inside windows of sigma rows
 * (stable), packed into slices of C rows; perm[k] is the original row of packed row k
 * (-1 for padding rows).  Slot j of packed row k of slice s is at
 * slice_ptr[s]*C + j*C + (k % C).  Padding entries hold value 0. */
typedef struct {
    int N, C, sigma, nslices, nnz_padded;
    int *slice_ptr;   /* nslices+1, widths prefix (in slots)            */
    int *cind;        /* nnz_padded                                      */
    int *perm;        /* nslices*C                                       */
    double *val64;    /* nnz_padded or NULL                              */
    float *val32;     /* nnz_padded or NULL                              */
} senk_sellcs;
/* Host packing (malloc'd arrays in *out; free with senk_sellcs_free_host). */
int senk_csr2sellcs(const double *val, const int *cind, const int *rptr, int N, int C, int sigma,
                    senk_sellcs *out);
void senk_sellcs_free_host(senk_sellcs *s);
/* Copy a host-packed SELL-C-sigma to the device (device pointers in *dev). */
int senk_sellcs_upload(senk_ctx ctx, const senk_sellcs *host, int want_f32, senk_sellcs *dev);
int senk_sellcs_free_device(senk_ctx ctx, senk_sellcs *dev);
int senk_spmv_sellcs_f64(senk_ctx ctx, const senk_sellcs *dev, const double *dx, double *dy);
int senk_spmv_sellcs_f32(senk_ctx ctx, const senk_sellcs *dev, const float *dx, float *dy);

/* SptrsvCsr_l / SptrsvCsr_u, ABMC form      senk_sparse.hpp:250-315.
 * cptr (host, cnum+1 entries) counts blocks of bsize rows.  L holds the strict lower
 * part; U holds the inverted diagonal first in each row ("L-DU", invDiag).  x may
 * alias y for the U solve. */
int senk_sptrsv_csr_l_abmc_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr,
                               const double *dx, double *dy, int N, const int *cptr, int cnum, int bsize);
int senk_sptrsv_csr_u_abmc_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr,
                               const double *dx, double *dy, int N, const int *cptr, int cnum, int bsize);
int senk_sptrsv_csr_l_abmc_f32(senk_ctx ctx, const float *dval, const int *dcind, const int *drptr,
                               const float *dx, float *dy, int N, const int *cptr, int cnum, int bsize);
int senk_sptrsv_csr_u_abmc_f32(senk_ctx ctx, const float *dval, const int *dcind, const int *drptr,
                               const float *dx, float *dy, int N, const int *cptr, int cnum, int bsize);
/* AMC form (senk_sparse.hpp:180-236): cptr counts rows. */
int senk_sptrsv_csr_l_amc_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr,
                              const double *dx, double *dy, int N, const int *cptr, int cnum);
int senk_sptrsv_csr_u_amc_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr,
                              const double *dx, double *dy, int N, const int *cptr, int cnum);
int senk_sptrsv_csr_l_amc_f32(senk_ctx ctx, const float *dval, const int *dcind, const int *drptr,
                              const float *dx, float *dy, int N, const int *cptr, int cnum);
int senk_sptrsv_csr_u_amc_f32(senk_ctx ctx, const float *dval, const int *dcind, const int *drptr,
                              const float *dx, float *dy, int N, const int *cptr, int cnum);

/* Sequential and block-Jacobi forms (senk_sparse.hpp:132-167, 328-379) need the
 * dependency levels of the factor, computed once on the host from its pattern:
 * rows of one level are independent, and every row's sum keeps the reference's CSR
 * order, so the result equals the sequential solve. */
typedef struct senk_trsv_plan_s *senk_trsv_plan;
/* upper = 0 for L (strict lower), 1 for U (diagonal first).  rptr/cind are host arrays. */
int senk_trsv_plan_create(senk_ctx ctx, const int *rptr, const int *cind, int N, int upper,
                          senk_trsv_plan *out);
int senk_trsv_plan_levels(senk_trsv_plan plan);
int senk_trsv_plan_destroy(senk_trsv_plan plan);
/* SptrsvCsr_l<T>(val,cind,rptr,x,y,N)       senk_sparse.hpp:132-143 */
int senk_sptrsv_csr_l_f64(senk_ctx ctx, senk_trsv_plan plan, const double *dval, const int *dcind,
                          const int *drptr, const double *dx, double *dy, int N);
/* SptrsvCsr_u<T>(val,cind,rptr,x,y,N)       senk_sparse.hpp:154-167 */
int senk_sptrsv_csr_u_f64(senk_ctx ctx, senk_trsv_plan plan, const double *dval, const int *dcind,
                          const int *drptr, const double *dx, double *dy, int N);
/* The same for T = float (senk_sparse.hpp:132-167 instantiated on float). */
int senk_sptrsv_csr_l_f32(senk_ctx ctx, senk_trsv_plan plan, const float *dval, const int *dcind,
                          const int *drptr, const float *dx, float *dy, int N);
int senk_sptrsv_csr_u_f32(senk_ctx ctx, senk_trsv_plan plan, const float *dval, const int *dcind,
                          const int *drptr, const float *dx, float *dy, int N);
/* SptrsvCsr_l / _u block-Jacobi form (bnum ranges of N/bnum rows; the plan is built from
 * the RemoveOffDiagonal factors, so ranges are independent).   senk_sparse.hpp:328-379 */
int senk_sptrsv_csr_l_bj_f64(senk_ctx ctx, senk_trsv_plan plan, const double *dval, const int *dcind,
                             const int *drptr, const double *dx, double *dy, int N, int bnum);
int senk_sptrsv_csr_u_bj_f64(senk_ctx ctx, senk_trsv_plan plan, const double *dval, const int *dcind,
                             const int *drptr, const double *dx, double *dy, int N, int bnum);
int senk_sptrsv_csr_l_bj_f32(senk_ctx ctx, senk_trsv_plan plan, const float *dval, const int *dcind,
                             const int *drptr, const float *dx, float *dy, int N, int bnum);
int senk_sptrsv_csr_u_bj_f32(senk_ctx ctx, senk_trsv_plan plan, const float *dval, const int *dcind,
                             const int *drptr, const float *dx, float *dy, int N, int bnum);
/* SptrsvBcsr_l / _u, ABMC form (bnl x bnw blocks)      senk_sparse.hpp:481-581 */
int senk_sptrsv_bcsr_l_abmc_f64(senk_ctx ctx, const double *dbval, const int *dbcind, const int *dbrptr,
                                const double *dx, double *dy, int N, const int *cptr, int cnum, int bsize,
                                int bnl, int bnw);
int senk_sptrsv_bcsr_u_abmc_f64(senk_ctx ctx, const double *dbval, const int *dbcind, const int *dbrptr,
                                const double *dx, double *dy, int N, const int *cptr, int cnum, int bsize,
                                int bnl, int bnw);
int senk_sptrsv_bcsr_l_abmc_f32(senk_ctx ctx, const float *dbval, const int *dbcind, const int *dbrptr,
                                const float *dx, float *dy, int N, const int *cptr, int cnum, int bsize,
                                int bnl, int bnw);
int senk_sptrsv_bcsr_u_abmc_f32(senk_ctx ctx, const float *dbval, const int *dbcind, const int *dbrptr,
                                const float *dx, float *dy, int N, const int *cptr, int cnum, int bsize,
                                int bnl, int bnw);
/* SptrsvBcsr_l / _u sequential form               senk_sparse.hpp:392-465 */
int senk_sptrsv_bcsr_l_f64(senk_ctx ctx, const double *dbval, const int *dbcind, const int *dbrptr,
                           const double *dx, double *dy, int N, int bnl, int bnw);
int senk_sptrsv_bcsr_u_f64(senk_ctx ctx, const double *dbval, const int *dbcind, const int *dbrptr,
                           const double *dx, double *dy, int N, int bnl, int bnw);
int senk_sptrsv_bcsr_l_f32(senk_ctx ctx, const float *dbval, const int *dbcind, const int *dbrptr,
                           const float *dx, float *dy, int N, int bnl, int bnw);
int senk_sptrsv_bcsr_u_f32(senk_ctx ctx, const float *dbval, const int *dbcind, const int *dbrptr,
                           const float *dx, float *dy, int N, int bnl, int bnw);

/* ------------------------------------------------------------------ BLAS-1 / BLAS-2
 * senk_blas1.hpp / senk_blas2.hpp.  Vectors are device pointers; scalars are host
 * values.  Dot/Nrm2 return the value on the host (synchronising); the _dev forms leave
 * it in device memory without synchronising.  Reductions accumulate in FP64 in a fixed
 * order, so repeated calls give identical results. */
int senk_copy_f64(senk_ctx ctx, const double *dx, double *dy, int N);                 /* :26-30  */
int senk_scal_f64(senk_ctx ctx, double a, double *dx, int N);                         /* :38-42  */
int senk_axpy_f64(senk_ctx ctx, double a, const double *dx, double *dy, int N);       /* :51-55  */
int senk_axpby_f64(senk_ctx ctx, double a, const double *dx, double b, double *dy, int N); /* :65-69 */
int senk_axpyz_f64(senk_ctx ctx, double a, const double *dx, const double *dy, double *dz, int N); /* :79-83 */
int senk_dot_f64(senk_ctx ctx, const double *dx, const double *dy, int N, double *result);   /* :92-98  */
int senk_nrm2_f64(senk_ctx ctx, const double *dx, int N, double *result);                    /* :106-112 */
int senk_dot_f64_dev(senk_ctx ctx, const double *dx, const double *dy, int N, double *dresult);
int senk_nrm2_f64_dev(senk_ctx ctx, const double *dx, int N, double *dresult);
int senk_hadprod_f64(senk_ctx ctx, const double *dx, double *dy, int N);              /* :120-124 */
int senk_haddiv_f64(senk_ctx ctx, const double *dx, double *dy, int N);               /* :132-136 */
int senk_copy_f32(senk_ctx ctx, const float *dx, float *dy, int N);
int senk_scal_f32(senk_ctx ctx, float a, float *dx, int N);
int senk_axpy_f32(senk_ctx ctx, float a, const float *dx, float *dy, int N);
int senk_axpby_f32(senk_ctx ctx, float a, const float *dx, float b, float *dy, int N);
int senk_axpyz_f32(senk_ctx ctx, float a, const float *dx, const float *dy, float *dz, int N);
int senk_hadprod_f32(senk_ctx ctx, const float *dx, float *dy, int N);
int senk_haddiv_f32(senk_ctx ctx, const float *dx, float *dy, int N);
int senk_dot_f32(senk_ctx ctx, const float *dx, const float *dy, int N, double *result);
int senk_nrm2_f32(senk_ctx ctx, const float *dx, int N, double *result);
/* Ggen / Grot are scalar host functions             senk_blas1.hpp:145-168 */
double senk_ggen(double a, double b, double *c, double *s);
void senk_grot(double c, double s, double *a, double *b);
/* Trsv (column-major upper triangular, leading dim n, order m)   senk_blas2.hpp:24-34.
 * Host arrays. */
void senk_trsv(const double *U, const double *b, double *x, int n, int m);
/* utils::Convert<double,float> on the device          senk_utils.hpp:103-110 */
int senk_convert_f64_f32(senk_ctx ctx, const double *din, float *dout, int n);
int senk_convert_f32_f64(senk_ctx ctx, const float *din, double *dout, int n);

/* test::RelativeResidualError (senk_test.hpp:11-22) on the device: *relres = ||b - A x|| /
 * ||b|| in FP64 (CSR A, device vectors); print != 0 also prints the reference's line
 * "# test %e". */
int senk_relative_residual_error(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr,
                                 const double *db, const double *dx, int N, int print, double *relres);
/* No reference counterpart: the same ||b - A x|| / ||b|| with no allocation at all, for
 * callers that cannot allocate (code under stream capture, tight serving loops).  The
 * result goes to device memory (*drelres) without synchronising; the partial sums live in
 * a 16 KiB array of the library's own on each device (looked up by senk_ctx_create), so
 * any N > 0 runs (at most 1024 workgroups, grid-stride over the rows), and calls that
 * overlap on one device (two streams or host threads) race on that array. */
int senk_relative_residual_error_noalloc(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr,
                                         const double *db, const double *dx, int N, double *drelres);

/* ------------------------------------------------------------------ matrix setup
 * senk_matrix.hpp.  Host arrays; integer outputs are bit-identical to the reference. */
int senk_check_structure(const int *cind, const int *rptr, int N);            /* :45-56, 1 = ok */
int senk_csr2csc(const double *val, const int *cind, const int *rptr,
                 double **cval, int **crind, int **ccptr, int N, int M);          /* :58-88 */
/* Csr2Sell (:90-142): returns the padded nnz in *nnz. */
int senk_csr2sell(const double *val, const int *cind, const int *rptr,
                  double **s_val, int **s_cind, int **s_len, int size, int N, int *nnz);
int senk_csr2bcsr(const double *val, const int *cind, const int *rptr,
                  double **bval, int **bcind, int **brptr, int N, int bnl, int bnw);   /* :144-211 */
int senk_bcsr2csr(const double *bval, const int *bcind, const int *brptr,
                  double **val, int **cind, int **rptr, int N, int bnl, int bnw, int *nnz); /* :213-241 */
int senk_padding(double **val, int **cind, int **rptr, int size, int *N);          /* :243-255 */
/* RemoveZeros (:29-43): entries equal to zero leave, in place; val/cind reallocated. */
int senk_remove_zeros(double **val, int **cind, int *rptr, int N);
/* Expand (:335-371): the full matrix from a stored triangle, key "L" (lower part with the
 * diagonal stored, as in a symmetric MatrixMarket file) or "U". */
int senk_expand(const double *tval, const int *tcind, const int *trptr,
                double **val, int **cind, int **rptr, int N, const char *key);
/* Duplicate (:373-389): a copy. */
int senk_duplicate(const double *tval, const int *tcind, const int *trptr,
                   double **val, int **cind, int **rptr, int N);

/* ------------------------------------------------------------------ MatrixMarket
 * io::ReadMatrixMarket (senk_io.hpp:41-103).  A real or integer coordinate file,
 * "general" (*shape = SENK_SHAPE_UNSYM) or "symmetric" (*shape = SENK_SHAPE_SYM: the
 * stored triangle is returned; senk_expand(..., "L") completes it).  removeZeros drops
 * every stored zero.  CSR output, rows in order and columns ascending, malloc'd.
 * SENK_ERR_IO if the file cannot be read; SENK_ERR_FORMAT for a banner other than
 * real/integer coordinate general/symmetric (pattern, complex, hermitian, skew-symmetric,
 * array), a bad size line, a short or out-of-range entry list. */
enum senk_shape { SENK_SHAPE_SYM = 0, SENK_SHAPE_UNSYM = 1 };
int senk_read_matrix_market(const char *filename, double **val, int **cind, int **rptr,
                            int *N, int *M, int *shape, int removeZeros);
/* The banner and size line only (senk_io.hpp:52-68: what the reader prints before the
 * entries): rows, columns, stored entries (before zero removal) and the shape. */
int senk_matrix_market_size(const char *filename, int *N, int *M, long long *entries, int *shape);
/* Split (:257-333): key "LD-U", "L-DU", "L-D-U" or "LU-D".  Unused outputs may be NULL. */
int senk_split(const double *val, const int *cind, const int *rptr,
               double **lval, int **lcind, int **lrptr, double **uval, int **ucind, int **urptr,
               double **diag, int N, const char *key, int invDiag);
int senk_get_diag(const double *val, const int *cind, const int *rptr, double **diag, int N); /* :391-404 */
int senk_scaling(double *val, const int *cind, const int *rptr, const double *ld, const double *rd, int N); /* :406-416 */
/* Scaling on the device. */
int senk_scaling_dev(senk_ctx ctx, double *dval, const int *dcind, const int *drptr,
                     const double *dld, const double *drd, int N);
/* AllocLevelZero (:551-650): ILU(p) symbolic fill; values kept, fill = 0. */
int senk_alloc_level_zero(double **val, int **cind, int **rptr, int N, int p);
/* AllocBlockZero (:652-666): bnl x bnw block closure of the pattern, zeros filled. */
int senk_alloc_block_zero(double **val, int **cind, int **rptr, int N, int bnl, int bnw);
/* RemoveOffDiagonal (:668-689) */
int senk_remove_off_diagonal(double **val, int **cind, int **rptr, int N, int bnum);
/* Reordering (:691-701, column relabel) and (:703-730, B = P A P^T). */
int senk_reordering_rp(double *val, int *cind, const int *rptr, const int *RP, int N);
int senk_reordering(double *val, int *cind, int *rptr, const int *LP, const int *RP, int N);

/* Ilu0 (:418-448) numeric factorization on the device, in place on the combined CSR
 * pattern.  The schedule is either ABMC colors (cptr in blocks of bsize rows, cnum > 0)
 * or, with cptr == NULL, the dependency levels of the pattern computed from the host
 * copies hrptr/hcind (pass NULL for those with ABMC colors).  Rows inside one ABMC block
 * run in order, and every entry receives its updates in increasing pivot column, so
 * the factor equals the reference's sequential Ilu0 (no contraction is used).  Returns
 * SENK_ERR_SCHEDULE if an entry couples two blocks of the same color. */
int senk_ilu0_f64(senk_ctx ctx, double *dval, const int *dcind, const int *drptr, int N,
                  const int *cptr, int cnum, int bsize, const int *hrptr, const int *hcind);
int senk_ilu0_f32(senk_ctx ctx, float *dval, const int *dcind, const int *drptr, int N,
                  const int *cptr, int cnum, int bsize, const int *hrptr, const int *hcind);

/* ------------------------------------------------------------------ graph ordering
 * senk_graph.hpp.  size_color, LP and RP are malloc'd like the reference. */
/* GetAMCPermutation (:250-277): size_color counts rows. */
int senk_amc_permutation(const int *cind, const int *rptr, int *num_color, int **size_color,
                         int **LP, int **RP, int N, int isSym);
/* GetABMCPermutation (:291-334): bmethod "simple" or "connect"; size_color counts blocks. */
int senk_abmc_permutation(const int *cind, const int *rptr, int *num_color, int **size_color,
                          int **LP, int **RP, int N, int bsize, int isSym, const char *bmethod);

/* ------------------------------------------------------------------ Krylov solvers */
typedef struct {
    int iters;               /* iteration count as the reference prints it               */
    int converged;           /* 1 when the stop test fired                               */
    int restarts;            /* restart cycles (GMRES/GCR) or refinement steps (IR)      */
    int inner_iters;         /* mixed precision: total inner iterations                  */
    int fell_back;           /* learned preconditioner: 1 if the monitor switched it off */
    int status;              /* SENK_OK or an error code                                  */
    double implicit_relres;  /* the residual estimate the stop test used (relative)      */
    double true_relres;      /* ||b - A x|| / ||b|| recomputed in FP64 after the solve   */
    double t_solve_ms;       /* device time of the solve (stream events)                 */
} senk_result;

/* Drop-in twins of the reference drivers.  Matrix, factors, b and x are device
 * pointers; cptr is a host array; nrm_b, outer/max_iter, m, N and epsilon keep the
 * reference's meaning.  x is the initial guess and the output.  When verbose != 0 the
 * lines the reference prints are printed (after the solve, from the device history).
 * hist (host, optional) receives the per-step residual the reference tests. */
int senk_gmresm_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr,
                    const double *db, double *dx, double nrm_b, int outer, int m, int N, double epsilon,
                    int verbose, senk_result *res, double *hist, int hist_cap);          /* senk_gmres.hpp:33-95 */
int senk_ilu_gmresm_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr,
                        const double *dlval, const int *dlcind, const int *dlrptr,
                        const double *duval, const int *ducind, const int *durptr,
                        senk_trsv_plan lplan, senk_trsv_plan uplan,
                        const double *db, double *dx, double nrm_b, int outer, int m, int N, double epsilon,
                        int verbose, senk_result *res, double *hist, int hist_cap);      /* :197-270 */
int senk_abmc_ilu_gmresm_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr,
                             const double *dlval, const int *dlcind, const int *dlrptr,
                             const double *duval, const int *ducind, const int *durptr,
                             const int *cptr, int cnum, int bsize,
                             const double *db, double *dx, double nrm_b, int outer, int m, int N, double epsilon,
                             int verbose, senk_result *res, double *hist, int hist_cap); /* :391-465 (AMC :293 with bsize 1) */
int senk_abmc_ilub_gmresm_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr,
                              const double *dblval, const int *dblcind, const int *dblrptr,
                              const double *dbuval, const int *dbucind, const int *dburptr,
                              const int *cptr, int cnum, int bsize,
                              const double *db, double *dx, double nrm_b, int outer, int m, int N, double epsilon,
                              int bnl, int bnw, int verbose, senk_result *res, double *hist, int hist_cap); /* :679-753 */
int senk_gcrm_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr,
                  const double *db, double *dx, double nrm_b, int outer, int m, int N, double epsilon,
                  int verbose, senk_result *res, double *hist, int hist_cap);            /* senk_gcr.hpp:30-72 */
int senk_abmc_ilu_gcrm_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr,
                           const double *dlval, const int *dlcind, const int *dlrptr,
                           const double *duval, const int *ducind, const int *durptr,
                           const int *cptr, int cnum, int bsize,
                           const double *db, double *dx, double nrm_b, int outer, int m, int N, double epsilon,
                           int verbose, senk_result *res, double *hist, int hist_cap);   /* :93-142 */
int senk_bicgstab_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr,
                      const double *db, double *dx, double nrm_b, int max_iter, int N, double epsilon,
                      int verbose, senk_result *res, double *hist, int hist_cap);        /* senk_bicgstab.hpp:29-87 */
int senk_abmc_ilu_bicgstab_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr,
                               const double *dlval, const int *dlcind, const int *dlrptr,
                               const double *duval, const int *ducind, const int *durptr,
                               const int *cptr, int cnum, int bsize,
                               const double *db, double *dx, double nrm_b, int max_iter, int N, double epsilon,
                               int verbose, senk_result *res, double *hist, int hist_cap); /* :107-176 */

/* Sequential, AMC and block-Jacobi forms.  lplan/uplan are senk_trsv_plan_create plans
 * of the factors' patterns (their dependency levels); the triangular solves give the
 * sequential result bit for bit.  ILUB forms take BCSR factors (bnl x bnw blocks) and
 * schedule their block rows by dependency levels. */
int senk_gmresm_bcsr_f64(senk_ctx ctx, const double *dbval, const int *dbcind, const int *dbrptr,
                         const double *db, double *dx, double nrm_b, int outer, int m, int N, double epsilon,
                         int bnl, int bnw, int verbose, senk_result *res, double *hist, int hist_cap); /* senk_gmres.hpp:113-175 */
int senk_amc_ilu_gmresm_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr,
                            const double *dlval, const int *dlcind, const int *dlrptr,
                            const double *duval, const int *ducind, const int *durptr,
                            const int *cptr, int cnum,
                            const double *db, double *dx, double nrm_b, int outer, int m, int N, double epsilon,
                            int verbose, senk_result *res, double *hist, int hist_cap);  /* senk_gmres.hpp:293-367 */
int senk_bj_ilu_gmresm_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr,
                           const double *dlval, const int *dlcind, const int *dlrptr,
                           const double *duval, const int *ducind, const int *durptr,
                           senk_trsv_plan lplan, senk_trsv_plan uplan, int bnum,
                           const double *db, double *dx, double nrm_b, int outer, int m, int N, double epsilon,
                           int verbose, senk_result *res, double *hist, int hist_cap);   /* senk_gmres.hpp:487-561 */
int senk_ilub_gmresm_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr,
                         const double *dblval, const int *dblcind, const int *dblrptr,
                         const double *dbuval, const int *dbucind, const int *dburptr,
                         const double *db, double *dx, double nrm_b, int outer, int m, int N, double epsilon,
                         int bnl, int bnw, int verbose, senk_result *res, double *hist, int hist_cap); /* senk_gmres.hpp:582-655 */
int senk_ilu_gcrm_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr,
                      const double *dlval, const int *dlcind, const int *dlrptr,
                      const double *duval, const int *ducind, const int *durptr,
                      senk_trsv_plan lplan, senk_trsv_plan uplan,
                      const double *db, double *dx, double nrm_b, int outer, int m, int N, double epsilon,
                      int verbose, senk_result *res, double *hist, int hist_cap);        /* senk_gcr.hpp:93-142 */
int senk_ilub_gcrm_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr,
                       const double *dblval, const int *dblcind, const int *dblrptr,
                       const double *dbuval, const int *dbucind, const int *dburptr,
                       const double *db, double *dx, double nrm_b, int outer, int m, int N, double epsilon,
                       int bnl, int bnw, int verbose, senk_result *res, double *hist, int hist_cap); /* senk_gcr.hpp:163-212 */
int senk_ilu_bicgstab_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr,
                          const double *dlval, const int *dlcind, const int *dlrptr,
                          const double *duval, const int *ducind, const int *durptr,
                          senk_trsv_plan lplan, senk_trsv_plan uplan,
                          const double *db, double *dx, double nrm_b, int max_iter, int N, double epsilon,
                          int verbose, senk_result *res, double *hist, int hist_cap);    /* senk_bicgstab.hpp:107-176 */
int senk_ilub_bicgstab_f64(senk_ctx ctx, const double *dval, const int *dcind, const int *drptr,
                           const double *dblval, const int *dblcind, const int *dblrptr,
                           const double *dbuval, const int *dbucind, const int *dburptr,
                           const double *db, double *dx, double nrm_b, int max_iter, int N, double epsilon,
                           int bnl, int bnw, int verbose, senk_result *res, double *hist, int hist_cap); /* senk_bicgstab.hpp:198-267 */

/* ------------------------------------------------------------------ handle API
 * Owned device matrix + preconditioner for repeated solves, mixed precision and the
 * learned correction (no reference counterparts). */
typedef struct senk_mat_s *senk_mat;
typedef struct senk_prec_s *senk_prec;

/* Upload a CSR matrix (host arrays).  flags: bit0 keep an FP32 copy of the values,
 * bit1 build SELL-C-sigma (C=64, sigma=1024) and use it for SpMV. */
int senk_mat_create(senk_ctx ctx, int N, const int *rptr, const int *cind, const double *val, int flags,
                    senk_mat *out);
int senk_mat_destroy(senk_mat A);
int senk_mat_spmv_f64(senk_mat A, const double *dx, double *dy);
int senk_mat_spmv_f32(senk_mat A, const float *dx, float *dy);
int senk_mat_nnz(senk_mat A, int *nnz);

/* SENK_IC0 / SENK_ICP: incomplete Cholesky (L D L^T) for a symmetric A (checked: pattern
 * and values; SENK_ERR_ARG otherwise), on A's pattern or its ILU(p) fill pattern (p).  The
 * reference announces IC(0)/IC(p) "using the symmetry" (README_ja.md:47) but has none.  The
 * numeric factorization is the device Ilu0 (for symmetric A the same L and pivots D); the
 * upper factor is then rewritten from the lower one, u_ij = d_i l_ji, so the
 * preconditioner L D L^T is exactly symmetric, as CG requires. */
enum senk_ilu_kind { SENK_ILU0 = 0, SENK_ILUP = 1, SENK_ILUB = 2, SENK_IC0 = 3, SENK_ICP = 4 };
enum senk_factor_precision { SENK_FACTOR_F64 = 0, SENK_FACTOR_F32 = 1, SENK_FACTOR_F16 = 2 };
typedef struct {
    int kind;       /* senk_ilu_kind                                                   */
    int p;          /* ILU(p) level                                                    */
    int bnl, bnw;   /* ILUB block shape                                                */
    int precision;  /* bitmask of senk_factor_precision copies kept: 1<<F64 | 1<<F32 ... */
    int csr_only;   /* 0: also keep the factors in the sliced solve layout and solve with
                       it (blocks that do not divide 64 as sub-blocks of gcd(bsize, 64)
                       rows); 1: CSR factors and kernels only.  Both give the same
                       bits where bsize divides 64; for other blocks the CSR kernel sums
                       each row across a wave and agrees to rounding.                   */
} senk_ilu_opts;
/* Factorize A (already ABMC-reordered with colors cptr/cnum/bsize) on the device:
 * host symbolic pattern (ILU(p)/ILUB), device numeric Ilu0 per color, device split
 * into L (strict lower) and U (inverted diagonal first). */
int senk_prec_create_ilu(senk_ctx ctx, senk_mat A, const int *rptr, const int *cind, const double *val,
                         const int *cptr, int cnum, int bsize, const senk_ilu_opts *opts, senk_prec *out);
int senk_prec_destroy(senk_prec M);
int senk_prec_apply_f64(senk_prec M, const double *dx, double *dy);
int senk_prec_apply_f32(senk_prec M, const float *dx, float *dy);
/* FP32 vectors, FP16 factor values (precision bit SENK_FACTOR_F16; no reference
 * counterpart, the reference keeps its factors in the solver's type T): each row
 * stores value * 2^-e in FP16 with one exponent e per row (the binary exponent of the
 * row's largest magnitude), the U inverted diagonal stays FP32, arithmetic is FP32. */
int senk_prec_apply_f16(senk_prec M, const float *dx, float *dy);
/* Triangular-solve launch mode of the sliced layout (no reference counterpart; the
 * reference runs one OpenMP parallel loop per color, senk_sparse.hpp:207-246):
 * 0 one launch per level; 1 one cooperative launch per triangle, each gather waiting
 * until the row it reads is final; 2 one launch per triangle of workgroups that draw
 * tickets in arrival order (no residency assumption), gathers waiting likewise; 3 the
 * same with one ticket stream per XCD, stream x taking the x-th eighth of every level
 * (workgroup b counts as XCD b mod 8; every stream has the same number of tickets per
 * level, so a wait is only on workgroups handed out earlier); the default, which -1
 * restores.  Results are
 * bit-identical in every mode.  Applications of one preconditioner must not overlap (one
 * stream). */
int senk_prec_set_trsv_mode(senk_prec M, int mode);
/* Split form of the mixed forms' inner BiCGStab (no reference counterpart; the
 * reference's BiCGStab is right-preconditioned, senk_bicgstab.hpp:129-161).  When the
 * factorization is an ILU(0) that changes no off-diagonal entry (no row i has entries k
 * < i and j != i with (k, j) in the pattern: a triangle-free graph such as a 5- or
 * 7-point stencil in any ordering), A = L D + U - E with D = diag(U), E = 2D - diag(A),
 * and the inner FP32 BiCGStab runs on L^-1 A U^-1 y = L^-1 r, each product being one U
 * and one L solve plus two vector passes and no SpMV (Eisenstat's trick); d = U^-1 y.
 * The inner FP32 CG (symmetric A) likewise runs as CG on L^-1 A L^-T preconditioned by
 * D^-1 with the same products.  On by default where it applies (also for IC(0) factors);
 * enable = 1 where it does not is SENK_ERR_UNSUPPORTED.  The FP64 solvers and the other
 * inner solvers are not affected. */
int senk_prec_set_split(senk_prec M, int enable);
/* Synchronise and report SENK_ERR_TIMEOUT if a single-launch solve since the last check
 * gave up a wait (never expected; senk_solve checks by itself). */
int senk_prec_check(senk_prec M);
int senk_prec_factor_nnz(senk_prec M, int *lnnz, int *unnz, int *nlevels_l, int *nlevels_u);
/* Copy the device factors back (host arrays sized by senk_prec_factor_nnz). */
int senk_prec_download(senk_prec M, int *lrptr, int *lcind, double *lval, int *urptr, int *ucind, double *uval);

/* SENK_CG: preconditioned conjugate gradients for symmetric positive definite systems
 * with a symmetric preconditioner (IC(0)/IC(p)); the reference announces CG
 * (README_ja.md:44) but has none.  Stop test and count as BiCGStab's; no learned
 * correction (SENK_ERR_ARG). */
enum senk_solver_kind { SENK_GMRES = 0, SENK_FGMRES = 1, SENK_GCR = 2, SENK_FGCR = 3, SENK_BICGSTAB = 4,
                        SENK_CG = 5 };
enum senk_ortho { SENK_MGS = 0, SENK_CGS2 = 1 };
typedef struct {
    int solver;          /* senk_solver_kind                                             */
    int m;               /* restart length                                               */
    int max_iter;        /* iteration budget (GMRES/GCR: rounded up to whole cycles)     */
    double eps;          /* relative residual target (relative to ||b||)                  */
    int ortho;           /* senk_ortho (GMRES family)                                    */
    int mixed;           /* 1: FP32 matrix/factors/inner Krylov + FP64 iterative refinement;
                            2: FP32 inner Krylov as the preconditioner of an FP64 flexible
                               GCR(outer_m) outer iteration (nested);
                            3: the FP64 solver with its ILU / ILUB application run on the
                               FP32 factor copies (FP64 vectors converted on the way in
                               and out; the preconditioner must keep an FP32 copy; FGMRES,
                               GCR, FGCR, BiCGStab or CG, not GMRES)                     */
    int factor_prec;     /* senk_factor_precision of the FP32 applications of the mixed
                            forms: F32 (default when the FP32 copy exists) or F16        */
    double inner_eps;    /* mixed 1, 2: inner relative tolerance per outer step; an outer
                            step that starts at relative residual rho uses
                            max(inner_eps, 0.5 * eps / rho), so the last inner solves
                            stop at what the outer tolerance still needs                 */
    int inner_max;       /* mixed: inner iteration budget per refinement step            */
    int check_every;     /* GCR: 1 = test ||r|| every step (reference: restarts only)    */
    int use_learned;     /* 1: add the learned correction: in every ILU application of a
                            flexible solver (FGCR, FGMRES, FP64 or as the mixed forms'
                            inner solver); with mixed != 0 and another inner solver, once
                            to each inner solve's result (the refinement's correction or
                            the outer flexible GCR's preconditioner application)         */
    int verbose;         /* print the reference's lines                                  */
    int outer_m;         /* mixed == 2: restart length of the FP64 outer GCR               */
} senk_solve_opts;
void senk_solve_opts_default(senk_solve_opts *o);
int senk_solve(senk_ctx ctx, senk_mat A, senk_prec M, const double *db, double *dx, const senk_solve_opts *opts,
               senk_result *res, double *hist, int hist_cap);

/* ------------------------------------------------------------------ learned correction
 * A message-passing network over the matrix graph that adds a correction to the ILU
 * application in flexible solves (no reference counterpart):
 *   s = |diag A|^-1/2, Ahat = S A S, x = [s r, z0/s]/rms(s r)   (z0 = ILU output)
 *   h1 = relu(x W1s + (Ahat x) W1n + b1), h2, h3 likewise (width 32)
 *   z = z0 + rms(s r) * s * (h3 wo + bo)
 * Layers 2 and 3 run on the matrix cores with BF16 weights and FP32 accumulation; the
 * parameter layout is W1s[2][32] W1n[2][32] b1[32] W2s[32][32] W2n[32][32] b2[32]
 * W3s[32][32] W3n[32][32] b3[32] wo[32] bo (SENK_GNN_PARAMS floats). */
typedef struct senk_gnn_s *senk_gnn;
enum { SENK_GNN_PARAMS = 4353 };
int senk_gnn_create(senk_ctx ctx, senk_gnn *out);   /* all parameters zero: no correction */
int senk_gnn_destroy(senk_gnn g);
int senk_gnn_set_params(senk_gnn g, const float *theta);   /* host arrays of SENK_GNN_PARAMS */
int senk_gnn_get_params(senk_gnn g, float *theta);
/* Bind the matrix the correction runs on (needs its FP32 values: senk_mat_create flag bit0).
 * Builds, on the host, the layers' tile order (breadth-first over tiles of 32 rows) and
 * their tiled copy of the graph (8 bytes per entry, widths padded to multiples of 8). */
int senk_gnn_bind(senk_gnn g, senk_mat A);
/* dz = dz + correction(dr, dz) on the bound matrix (dz holds the ILU output on entry). */
int senk_gnn_apply_f32(senk_gnn g, const float *dr, float *dz);
/* Mean training loss ||A z - r||^2/||r||^2 over nvec residuals (device FP32, nvec rows of N)
 * with M's FP32 factors, and its gradient (host, SENK_GNN_PARAMS doubles, may be NULL). */
int senk_gnn_loss_grad(senk_gnn g, senk_mat A, senk_prec M, const float *dR, int nvec, double *loss,
                       double *grad);
/* The same with z0 for row k of dR taken from row k of dZ (device FP32) instead of M's
 * ILU application: the correction of another preconditioner's result. */
int senk_gnn_loss_grad_pairs(senk_gnn g, senk_mat A, senk_prec M, const float *dR, const float *dZ, int nvec,
                             double *loss, double *grad);
/* Adam on the device over the residual sets of nmat matrices (dR[k]: nvec[k] FP32 rows of
 * A[k]'s N); loss_hist (host, epochs) receives the mean loss of each epoch. */
int senk_gnn_train(senk_gnn g, int nmat, const senk_mat *A, const senk_prec *M, const float *const *dR,
                   const int *nvec, int epochs, int batch, double lr, unsigned seed, double *loss_hist,
                   int verbose);
/* The same with z0 rows from dZ[k] (as senk_gnn_loss_grad_pairs; dZ == NULL: the ILU). */
int senk_gnn_train_pairs(senk_gnn g, int nmat, const senk_mat *A, const senk_prec *M, const float *const *dR,
                         const float *const *dZ, const int *nvec, int epochs, int batch, double lr,
                         unsigned seed, double *loss_hist, int verbose);
/* Residual/error pairs from the plain FP64 ILU-GMRES(m) of the handle API: after each of
 * up to ncycles restart cycles from x0 = 0, row k of dR receives r = b - A x and, when
 * dxstar is given, row k of dE receives x* - x (device FP64).  *count = rows written. */
int senk_gnn_collect(senk_ctx ctx, senk_mat A, senk_prec M, const double *db, const double *dxstar, int m,
                     int ncycles, double *dR, double *dE, int *count);
/* Training pairs for the correction of the nested form's inner solve (senk_solve with
 * opts->mixed = 2 and an inner solver other than FGCR / FGMRES): runs that solve from
 * x0 = 0 without a correction and, at every outer step, stores the normalised FP32 residual
 * the inner solve received as row k of dR and its result as row k of dZ (device FP32, cap
 * rows of N each); *count = rows written. */
int senk_gnn_collect_nested(senk_ctx ctx, senk_mat A, senk_prec M, const double *db,
                            const senk_solve_opts *opts, float *dR, float *dZ, int cap, int *count);
/* Use g (bound to M's matrix) as M's learned correction in solves with use_learned = 1;
 * g == NULL detaches.  A solve whose FP64 residual stagnates switches the correction off
 * for the rest of that solve (senk_result.fell_back). */
int senk_prec_attach_gnn(senk_prec M, senk_gnn g);

/* ------------------------------------------------------------------ synthetic matrices
 * Generators for the benchmark systems (no reference counterpart).  Outputs are
 * malloc'd sorted CSR arrays.  Every choice a generator makes is stated in its source. */
int senk_gen_poisson7(int nx, int ny, int nz, int *N, int **rptr, int **cind, double **val);
int senk_gen_convdiff27(int n, double peclet, int *N, int **rptr, int **cind, double **val);
int senk_gen_vardiff7(int n, int seed, double corr_len, double sigma, int *N, int **rptr, int **cind, double **val);
int senk_gen_unstructured(int n, int seed, int *N, int **rptr, int **cind, double **val);
int senk_gen_elasticity3(int n, int seed, double nu, int *N, int **rptr, int **cind, double **val);
/* x* ~ U[-1,1] from std::mt19937(seed) + uniform_real_distribution<double>(-1,1) in
 * index order (main.cpp:117-119). */
int senk_gen_uniform(int seed, int n, double *out);

#ifdef __cplusplus
}
#endif
#endif
