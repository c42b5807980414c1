/*
 * pdcs_oracle.h -- CPU FP64 restatement of the PDCS PDHG hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  The HIP library (libpdcs_hip.so) is tested against
 * this oracle.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  It is never linked into, called by, or used
 * as a fallback for the product path.
 *
 * Problem form (reference README.md:1-40, src/pdcs_gpu/rpdhg_alg_gpu_gen.jl:1219):
 *     min c'x   s.t.  G x - h in K,   bl <= x[0:nb] <= bu
 * K = {0}^mzero x R+^mnonneg x SOC(q_1) x ... x Kexp^nexp x Kexp*^ndexp, rows in
 * exactly that order (recover_soc_cone_rotated_exp_cone_indices,
 * rpdhg_alg_gpu_gen.jl:60-118).  G is CSR, 0-based, int32 indices.
 */
#ifndef PDCS_ORACLE_H
#define PDCS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int64_t n, m, nb;
    const int32_t *rowptr;   /* m+1 */
    const int32_t *colidx;   /* nnz */
    const double *vals;      /* nnz */
    const double *c, *h;     /* n, m */
    const double *bl, *bu;   /* nb */
    int64_t mzero, mnonneg;
    int64_t nsoc;
    const int64_t *soc;      /* nsoc cone dimensions */
    int64_t nexp, ndexp;
} oracle_problem;

typedef struct {
    int64_t max_outer_iter;          /* 10000  rpdhg_alg_gpu_gen.jl:415 */
    int64_t max_inner_iter;          /* 500000 :416 */
    double abs_tol, rel_tol;         /* 1e-6   :417-418 */
    double time_limit;               /* inf    :426 */
    int64_t check_terminate_freq;    /* 2000   :423 */
    int64_t duality_gap_restart_freq;/* 2000   :433 */
    int64_t kkt_restart_freq;        /* 2000   :431 */
    int32_t use_kkt_restart;         /* 0      :430 */
    int32_t use_duality_gap_restart; /* 1      :432 */
    int32_t use_preconditioner;      /* 1      :409 */
    int32_t ruiz_iters;              /* 10     rpdhg_alg_gpu_gen.jl:937-1004 */
    int32_t pock_chambolle;          /* 1 (alpha = 1) */
    int32_t warm_start;              /* 0 */
    int32_t verbose;                 /* 0 */
    /* bisection overshoot semantics of binary_search_duality_gap_diagonal!
     * (rpdhg_alg_gpu_gen_scaling.jl:86-92): 0 = as the reference's Julia parses
     * `tRight = t, tLeft = t / 2, break` (plain break, bracket unchanged);
     * 1 = the evident intent (tRight = t; tLeft = t/2). */
    int32_t bisect_intent;
    /* termination ratio: 0 = README l_inf primal denominator
     * 1+max(||h||,||Gx||,||proj_K(Gx-h)||); 1 = code variant
     * 1+max(||proj_Lambda(c-G'y)||,||h||,||Gx||) (rpdhg_alg_gpu_gen_scaling.jl:153-162). */
    int32_t code_primal_denominator;
    /* outer iteration at which the pessimistic re-solve
     * (rpdhg_alg_gpu_gen_scaling.jl:511-638) starts whatever its triggers say;
     * 0 = only on the triggers (rpdhg_alg_gpu_gen_scaling.jl:811-836).  Not a
     * parameter of the reference: it makes the re-solve reachable in tests. */
    int64_t pessimistic_at;
} oracle_params;

typedef struct {
    int32_t exit_code;     /* 0 optimal, 1 max_iter, 6 time_limit, 8 nan, 9 inf (termination.jl:181-260) */
    int32_t sequence;      /* 0 = current iterate met the test, 1 = averaged iterate */
    int64_t iter;          /* PDHG attempts (accepted + rejected) */
    int64_t restarts;
    double pobj, dobj;
    double rel_primal_res, rel_dual_res, rel_gap;           /* README l_inf */
    double abs_primal_res, abs_dual_res, abs_gap;
    double l2_rel_primal_res, l2_rel_dual_res;
    double rel_primal_res_code;                             /* code-variant denominator */
    double omega;
    double time_solve;
    int32_t pessimistic_flag;  /* 1 if the pessimistic re-solve ran */
    double pessimistic_norm;   /* its power-method estimate of ||G||_2 (scaled G), 0 if it did not run */
    int64_t pessimistic_iter;  /* attempts when it started */
} oracle_result;

void oracle_default_params(oracle_params *p);

/* whole solve; x_out (n) and y_out (m) receive the unscaled primal/dual of the
 * sequence that met the test (may be NULL).  x0/y0 are warm starts when
 * p->warm_start (unscaled). Returns 0 or a negative error code. */
int oracle_solve(const oracle_problem *prob, const oracle_params *p,
                 const double *x0, const double *y0,
                 double *x_out, double *y_out, oracle_result *res);

/* run only the first `attempts` PDHG attempts (CPU baseline sample); returns elapsed seconds */
double oracle_time_attempts(const oracle_problem *prob, int64_t attempts, int64_t *done);

/* -------- kernel-level pieces (exposed for tests) -------- */
/* y <- G x  (CSR) */
void oracle_spmv(int64_t m, const int32_t *rowptr, const int32_t *colidx,
                 const double *vals, const double *x, double *y);
/* in-place projections */
void oracle_proj_soc(double *v, int64_t len);
/* scaled SOC {(t,z): ||d .* z|| <= t}; d[0] unused, d2 = d.^2; *xi_ws warm start */
void oracle_proj_soc_scaled(double *v, const double *d, const double *d2, int64_t len,
                            double *xi_ws, double abs_tol, double rel_tol, double *tmp);
/* projection onto the dual cone of K in scaled space (codes 1/4/6/12/15), in place */
void oracle_proj_dual(const oracle_problem *prob, double *y, const double *dsc,
                      const double *dsc2, double *warm, double abs_tol, double rel_tol);
/* projection onto K (codes 2/3/5/13/16), in place */
void oracle_proj_primal_cone(const oracle_problem *prob, double *r);
/* exponential cone K_exp = cl{(r,s,t): s>0, s exp(r/s) <= t} and its dual */
void oracle_proj_exp(double *v);
void oracle_proj_dual_exp(double *v);

/* Ruiz(10) + Pock-Chambolle(alpha=1) scaling; outputs Dr (n), Dl (m) products
 * and the scaled CSR values (nnz) */
void oracle_scale(const oracle_problem *prob, int ruiz_iters, int pock, double *Dr,
                  double *Dl, double *vals_scaled);

/* Forward mode by dual numbers: constant-step PDHG on the original data started
 * at (x, y) with data tangents (dc, dh, dG; NULL = 0) for at most `iters`
 * iterations (stops when the change of values and tangents is below tol);
 * x, y are updated in place (a refinement of the solution) and dx, dy receive
 * the tangents.  Zero, nonnegative and SOC rows only (returns -2 otherwise). */
int oracle_tangent(const oracle_problem *P, double *x, double *y, const double *dc, const double *dh,
                   const double *dG, int64_t iters, double tol, double *dx, double *dy, int64_t *done);

/* README residuals of (x, y) in the original space */
void oracle_kkt(const oracle_problem *prob, const double *x, const double *y,
                oracle_result *res);

#ifdef __cplusplus
}
#endif
#endif
