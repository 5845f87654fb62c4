// Internal interface of the device-resident Krylov drivers.
#pragma once

#include "internal.hpp"

struct senk_gnn_s;

namespace senk_amd {

struct MatView {
    int N = 0, nnz = 0;
    const int *rptr = nullptr, *cind = nullptr;
    const double *v64 = nullptr;
    const float *v32 = nullptr;
    double avg_row = 0;
    // optional SELL-C-sigma copy (device); per value type, the order in which its slices
    // are handed to the workgroups (null: slice order) and whether the matrix stream is
    // read with non-temporal loads (both chosen by timing when a colored preconditioner
    // is built on the matrix; neither changes a result bit)
    const senk_sellcs *sell = nullptr;
    const int *sell_order64 = nullptr, *sell_order32 = nullptr;
    bool sell_nt64 = false, sell_nt32 = false;
    // optional BCSR form used for SpMV (device); bnl == 0 when absent
    int bnl = 0, bnw = 0;
    const int *brptr = nullptr, *bcind = nullptr;
    const double *bv64 = nullptr;
    const float *bv32 = nullptr;
};

// The learned correction as a preconditioner component (gnn.hip); off = switched off
// by a solver's stagnation monitor, leaving plain ILU.
struct LearnedPrec {
    senk_gnn_s *net = nullptr;
    bool off = false;
};

struct PrecView {
    int N = 0;
    int kind = 0;  // 0 none, 1 ILU (CSR factors), 2 ILUB (BCSR factors)
    Schedule ls, us;
    Schedule lts, uts;  // the schedules of the solve layouts lt, ut (sub-blocks where ls / us
                        // have blocks a wave cannot hold whole)
    // CSR factors: L strict lower, U with the inverted diagonal first
    const int *lrptr = nullptr, *lcind = nullptr, *urptr = nullptr, *ucind = nullptr;
    const double *l64 = nullptr, *u64 = nullptr;
    const float *l32 = nullptr, *u32 = nullptr;
    // optional solve layouts of the CSR factors (used when they hold the precision)
    TriLayout *lt = nullptr, *ut = nullptr;  // the ticketed sweep advances their counters
    // FP32 applications read the layouts' FP16 factor copies (per-row scaled)
    bool f16 = false;
    // BCSR factors
    int bnl = 0, bnw = 0;
    const int *blrptr = nullptr, *blcind = nullptr, *burptr = nullptr, *bucind = nullptr;
    const double *bl64 = nullptr, *bu64 = nullptr;
    const float *bl32 = nullptr, *bu32 = nullptr;
    // D-ILU (an ILU(0) whose off-diagonal factor entries are A's own: A = L D + U - E with
    // D = diag(U), E = 2 D - diag(A)): D and E in FP32, matrix order, for the split form
    // of the mixed forms' inner BiCGStab; null when the factorization is not D-ILU or the
    // split form is switched off
    const float *dilu_d = nullptr, *dilu_e = nullptr;
    // A (to a few ulps) and the factorization symmetric: the split form of the inner CG applies
    bool symmetric = false;
    // FP64 solves: the ILU / ILUB application in FP32 (the FP32 factor copies) between
    // conversions through in32 / out32 (N floats each)
    bool apply32 = false;
    float *in32 = nullptr, *out32 = nullptr;
    // learned correction added to the ILU application (flexible solvers only)
    LearnedPrec *learned = nullptr;
    int *learned_off = nullptr;  // device flag: 1 = correction disabled by the monitor
};

template <typename T>
int spmv(Ctx *c, const MatView &A, const T *x, T *y, const int *stop = nullptr);
template <typename T>
int prec_apply(Ctx *c, const PrecView &M, const T *in, const double *in_scale, T *out, T *tmp, const int *stop);

struct SolveParams {
    int solver = SENK_GMRES;
    int m = 30;
    int outer = 1;       // GMRES/GCR cycles
    int max_iter = 1000; // BiCGStab
    double eps = 1e-8;
    double nrm_b = 1;
    int ortho = SENK_MGS;
    int check_every = 0;
    int verbose = 0;
    int unprec_direct = 0;  // GMRES with no preconditioner: x += sum y_k v_k (senk_gmres.hpp:80-82)
    int split = 0;          // BiCGStab from x = 0: the split form when M has D-ILU vectors
};

// x (device) is the initial guess and the output.  hist (host) optional.
template <typename T>
int krylov_solve(Ctx *c, const MatView &A, const PrecView &M, const T *b, T *x, const SolveParams &p,
                 senk_result *res, double *hist, int hist_cap);

int ir_solve(Ctx *c, const MatView &A, const PrecView &M, const double *b, double *x, const senk_solve_opts &o,
             double nrm_b, senk_result *res, double *hist, int hist_cap);

// Rows recorded at every outer step of the nested solve: the normalised FP32 residual the
// inner solve received and the inner solve's result (before any learned correction).
struct NestedRecord {
    float *R = nullptr, *Z = nullptr;  // device, cap rows of N
    int cap = 0, count = 0;
};
// FP64 flexible GCR(o.outer_m) whose preconditioner is an FP32 inner Krylov solve.
int nested_solve(Ctx *c, const MatView &A, const PrecView &M, const double *b, double *x,
                 const senk_solve_opts &o, double nrm_b, senk_result *res, double *hist, int hist_cap,
                 NestedRecord *rec = nullptr);

// Learned correction (gnn.hip).
int gnn_create(Ctx *c, senk_gnn_s **out);
void gnn_destroy(senk_gnn_s *net);
int gnn_set_params(senk_gnn_s *net, const float *host_theta);
int gnn_get_params(senk_gnn_s *net, float *host_theta);
int gnn_bind(senk_gnn_s *net, int N, const int *drptr, const int *dcind, const float *dval);
LearnedPrec *gnn_learned(senk_gnn_s *net);
int gnn_apply_f32(senk_gnn_s *net, const float *r, float *z);
// Z (optional): row k holds z0 for row k of R (else z0 = M's FP32 ILU application).
int gnn_loss_grad(senk_gnn_s *net, const MatView &A, const PrecView &M, const float *R, const float *Z, int nvec,
                  double *loss, double *grad_host);
int gnn_train(senk_gnn_s *net, int nmat, const MatView *As, const PrecView *Ms, const float *const *Rs,
              const float *const *Zs, const int *nvecs, int epochs, int batch, double lr, unsigned seed,
              double *loss_hist, int verbose);

// Restart residuals (and errors against xstar) of the plain FP64 GMRES(m) with M.
int collect_pairs(Ctx *c, const MatView &A, const PrecView &M, const double *b, const double *xstar, int m,
                  int ncycles, double *R, double *E, int *count);

// ||b - A x|| / nrm_b in FP64 (synchronises).
int true_relres(Ctx *c, const MatView &A, const double *b, const double *x, double nrm_b, double *out);

// noalloc.hip: the address of the file-scope scratch array on `device` (current on entry),
// looked up once per device; ||b - A x|| / ||b|| into device memory with that scratch.
int noalloc_init(int device);
int relres_noalloc(Ctx *c, const double *dval, const int *dcind, const int *drptr, const double *db,
                   const double *dx, int N, double *drelres);

}  // namespace senk_amd
