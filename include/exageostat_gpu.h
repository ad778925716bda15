/*
 * exageostat_gpu.h -- C ABI of the MI355X-native exact-dense Gaussian
 * likelihood path.
 *
 * The entry points keep the names, argument meaning and return values of the
 * reference's C API for this path, so a caller (NLOPT objective, example
 * driver, R wrapper) links against libexageostat_gpu.so instead of
 * libexageostat + Chameleon + StarPU.  Each declaration cites the reference
 * interface it replaces.  The MORSE_* spellings are the legacy names of the
 * same functions (Chameleon was called MORSE).
 *
 * Deviations (deliberate):
 *  - no exit(): a non-SPD covariance or an unknown kernel string makes the
 *    likelihood return -HUGE_VAL and sets MLE_data.status and
 *    exageostat_gpu_last_status(); the reference calls exit()
 *    (misc/include/MLE_misc.h:343-348, codelet_zcmg.c:306-310);
 *  - descriptors are this library's own device tile matrices: the struct is
 *    source-compatible with the fields the hot path reads (m, n, mb, nb, mt,
 *    nt, mat, dtyp), not binary-compatible with Chameleon's CHAM_desc_t;
 *  - MLE_data has every field of the reference struct, in its order
 *    (misc/include/MLE_misc.h:184-328), plus fields of this implementation
 *    at the end.
 */
#ifndef EXAGEOSTAT_GPU_H
#define EXAGEOSTAT_GPU_H

#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* misc/include/MLE_misc.h:111-115 (z == NULL: 2-D locations). */
typedef struct {
  double *x;
  double *y;
  double *z;
} location;

/* Chameleon enumerations used by the path. */
typedef int CHAM_enum;
#define ChamUpper 121
#define ChamLower 122
#define ChamUpperLower 123
#define ChamRealFloat 2
#define ChamRealDouble 3
#define ChamLeft 141
#define ChamRight 142
#define ChamNoTrans 111
#define ChamTrans 112
#define ChamNonUnit 131
#define ChamUnit 132

/* Device tile matrix (replaces CHAM_desc_t for this path). */
typedef struct exageostat_desc_s {
  void *mat;   /* engine object owning the device tiles                      */
  int dtyp;    /* ChamRealDouble (mixed maps are per tile, see MLE_data)     */
  int mb, nb;  /* tile size                                                  */
  int bsiz;    /* mb * nb                                                    */
  int lm, ln;  /* padded size = mt*mb, nt*nb                                 */
  int m, n;    /* logical size                                               */
  int mt, nt;  /* tile counts                                                */
  int uplo;    /* ChamLower: only tiles m >= n stored; ChamUpperLower: all   */
} CHAM_desc_t;

/* Replaces RUNTIME_sequence_t / RUNTIME_request_t: an ordered HIP stream. */
typedef struct exageostat_sequence_s {
  int status;
  void *impl;
} RUNTIME_sequence_t;
typedef struct exageostat_request_s {
  int status;
} RUNTIME_request_t;

/* misc/include/MLE_misc.h:130-178: the summary a driver fills and prints.
 * The library writes final_loglik and estimated_theta (when the caller has
 * allocated it) after every likelihood evaluation, and the prediction, MLOE
 * and Fisher fields after those calls, as the reference does. */
typedef struct {
  int problem_size;
  char *computation;
  char *kernel;
  char *ds_type;
  char *precision;
  int z_sample;
  int dense_ts;
  int lr_ts;
  int lr_acc;
  int lr_maxrank;
  int ncores;
  int ngpus;
  int p;
  int q;
  int num_params;
  double *initial_theta;
  double *starting_theta;
  double *estimated_theta;
  double final_loglik;
  double time_per_iteration;
  double cholesky_per_iter;
  double flops_per_iteration;
  double generation_per_iter;
  double solve_per_iter;
  double total_mle_time;
  double avg_exec_time_gen_stage;
  double avg_flops_gen_stage;
  double mse_pred1;
  double mse_pred2;
  double mse_pred3;
  double mse_pred;
  double trace_pred_sum;
  double trace_pred_mean;
  double total_pred_time;
  double total_pred_flops;
  double mloe;
  double mmom;
  char *mloe_exec;
  double total_mloe_mmom_time;
  double matrix_gen_mloe_mmom_time;
  double cho_fact_mloe_mmom_time;
  double loop_mloe_mmom_time;
  double total_mloe_mmom_flops;
  double fisher_00;
  double fisher_11;
  double fisher_22;
  double fisher_time;
} output;

/* misc/include/MLE_misc.h:350 (defined by the library; a driver that defines
 * its own `output results;`, as examples/examples.c:22 does, takes
 * precedence). */
extern output results;

/* MLE_data: every field of the reference struct in the reference's order
 * (misc/include/MLE_misc.h:184-328), so the reference's drivers and their
 * argument parsing (examples/examples.c) compile unchanged; fields of this
 * implementation follow at the end.  Fields the exact-dense GPU path does
 * not use (HiCMA, out-of-core, non-Gaussian, file paths of real datasets
 * other than through readLocsFile/readObsFile) are accepted and ignored. */
typedef struct {
  double variance;
  double variance1;
  double variance2;
  double variance3;
  char *computation;    /* "exact" or "diag_approx"                          */
  char *c_fun;
  int test;
  int async;            /* 1: the *_Async entry points (same results here)  */
  int iter_count;
  location l1;          /* locations of the observations                    */
  location lmiss;       /* prediction: held-out locations                   */
  location lobs;        /* prediction: observed locations                   */
  location lm;
  void *descC;          /* Sigma, lower tiles                               */
  void *descsubC11;
  void *descsubC12;
  void *descsubC21;
  void *descsubC22;
  void *descZ;
  void *descZ1;
  void *descZ2;
  void *descZ3;
  double *Adense;
  double *Adense2;
  void *descZcpy;
  void *descdet;
  void *descsum;
  void *descproduct;
  void *descproduct1;
  void *descproduct2;
  void *descproduct3;
  void *descZmiss;
  void *descC11;
  void *descC21;
  void *descC12;
  void *descC22;
  void *descZactual;
  void *descZobs;
  void *descmse;
  void *descmse1;
  void *descmse2;
  void *descmse3;
  void *descZtrace;
  void *sequence;
  void *request;
  int verbose;
  int check;
  int log;              /* 1: iteration lines also go to pFileLog            */
  double avg_exec_time_per_iter;
  double total_exec_time;
  double avg_flops_per_iter;
  double avg_generation_per_iter;
  double avg_solve_per_iter;
  double avg_cholesky_per_iter;
  double avg_exec_time_gen_stage;
  double avg_flops_gen_stage;
  double final_loglik;
  char *locsFPath;
  char *timeFPath;
  char *obsFPath;
  char *obsFPath2;
  char *obsFPath3;
  char *actualZFPath;
  char *actualZFPath2;
  char *actualZFPath3;
  char *actualZLocFPath;
  char *actualTimeFPath;
  double det;           /* sum of log L_ii (logdet = 2*det)                  */
  double sum;
  double dotp;          /* ||L^-1 z||^2                                      */
  double dotp1;
  double dotp2;
  double dotp3;
  float sdotp;          /* precision 1: the dot product in FP32              */
  double mserror;
  double mserror1;
  double mserror2;
  double mserror3;
  char *dm;             /* "ed" Euclidean, "gc" great circle                 */
  int diag_thick;       /* diag_approx: tile band factored; mixed: FP64 band */
  char *nFileLog;
  FILE *pFileLog;
  int hicma_maxrank;
  int hicma_data_type;
  void *hicma_descC;
  void *hicma_descZ;
  void *hicma_descCD;
  void *hicma_descCUV;
  void *hicma_descCrk;
  void *hicma_descZcpy;
  void *hicma_descdet;
  void *hicma_descproduct;
  void *hicma_descC12D;
  void *hicma_descC12UV;
  void *hicma_descC12rk;
  void *hicma_descC22D;
  void *hicma_descC22UV;
  void *hicma_descC22rk;
  double hicma_acc;
  void *hsequence;
  void *hrequest;
  int opt_tol;
  int opt_max_iters;
  int ooc;
  char *kernel_fun;     /* "univariate_matern_stationary", ...              */
  int precision;        /* 0 double, 1 single, 2 mixed (FP64/FP32/BF16 tiles) */
  void *desctemp;
  void *desctemp22;
  void *desck_t;
  void *desck_a;
  void *desck_ttmp;
  void *desck_atmp;
  void *descK_t;
  void *descK_a;
  void *descexpr1;
  void *descexpr2;
  void *descexpr3;
  void *descexpr4;
  void *descestimatedalpha;
  void *desctruthalpha;
  void *desc_mloe_mmom;
  void *descmloe;
  void *descmmom;
  double expr1;
  double expr2;
  double expr3;
  double expr4;
  double mloe;          /* MLOE of the last EXAGEOSTAT_dmle_mloe_mmom_Tile     */
  double mmom;          /* MMOM of the last EXAGEOSTAT_dmle_mloe_mmom_Tile     */
  int mloe_mmom;
  int mloe_mmom_async;
  int mspe;
  char *recovery_file;  /* "": none; else replayed / appended per evaluation */
  char *checkpoint_file;
  int time_slots;
  int idw;
  int fisher;
  void *descr;
  void *descrcpy;
  int seed;
  /* ---- fields of this implementation ---- */
  void *gpu;            /* engine context (EXAGEOSTAT_dmle_Call / MLE_Finalize) */
  int silent;           /* 1: no per-iteration line on stderr                */
  int status;           /* 0 ok; >0 first non-positive pivot + 1; <0 error   */
  double logdet;        /* last log|Sigma|                                   */
  double time_gen;      /* last evaluation, seconds per phase                */
  double time_facto;
  double time_logdet;
  double time_solve;
  double time_total;
  double mp_fp32_corr;  /* mixed: off-diagonal tiles whose largest
                           correlation is >= this stay FP32, others BF16     */
  int mp_bf16;          /* mixed: 1 enables the BF16 tier                    */
  int mp_ir_steps;      /* mixed: at most this many FP64 refinement steps    */
  double mp_fp64_corr;  /* mixed: off-diagonal tiles whose largest
                           correlation is >= this stay FP64                  */
  int mp_lolo;          /* mixed: 1 adds the lo*lo term to the split-BF16
                           products of the FP32 tiles (four products instead
                           of three: smaller log-det bias, ~6% slower)       */
  double mp_ir_tol;     /* mixed: > 0 stops the refinement once a step
                           changes z^T Sigma^-1 z by <= mp_ir_tol (relative),
                           or once the changes still to come, extrapolated
                           from the last two steps (linear convergence,
                           ratio < 1/2), sum to <= mp_ir_tol;
                           if mp_ir_steps steps do not get there, status is
                           -3 (the likelihood is still returned)             */
  int mp_ir_done;       /* mixed, output: refinement steps run               */
  double mp_ir_change;  /* mixed, output: relative change of the last step   */
  int mp_lean;          /* read by EXAGEOSTAT_dmle_Call: 1 allocates no FP64
                           covariance.  Only the mixed (precision 2) and
                           single (precision 1) likelihoods, prediction and
                           MLOE/MMOM run; the refinement regenerates Sigma's
                           tiles in FP64 for its products instead of reading
                           them, so the device holds the mixed factor only
                           (about half of the FP64 matrix's bytes).  MLE_zvg,
                           the FP64 likelihood, the gradient, the Fisher
                           matrix and descC fail with status -1.            */
  int mp_grad_f32;      /* mixed, gradient request: 1 runs the Takahashi
                           sweep's column stage (Sigma^-1 from the mixed
                           factor) on the FP32 matrix cores, about 0.6 of
                           its FP64 time.  Its rounding error grows with the
                           condition number of Sigma: 4e-8 of the largest
                           gradient component at 40,000 locations with a
                           0.01 nugget, 3.7e-3 (1.8e-4 with 0) for the
                           bivariate kernel at 50,000 locations.  2 does the
                           same and, when mp_grad_check exceeds 1e-6,
                           recomputes the gradient with the FP64 sweep.
                           0 (the default) runs the sweep in FP64.          */
  double mp_grad_check; /* mixed, output of mp_grad_f32 1 and 2:
                           ||Z z - x0|| / ||x0||, Z the FP32-swept inverse
                           and x0 the solve with the mixed factor; 0 when
                           the FP32 sweep did not run                       */
} MLE_data;

/* ---------------------------------------------------------------- setup -- */
/* src/compute/MLE.c:233-260 (CHAMELEON_Init(ncores, gpus)). */
void exageostat_init(int *ncores, int *gpus, int *dts, int *lts);
/* src/compute/MLE.c (CHAMELEON_Finalize). */
void exageostat_finalize(void);
/* exageostat_exact/src/compute/MLE_dexact.c:1938-2017. */
void EXAGEOSTAT_dmle_Call(MLE_data *data, int ncores, int gpus, int dts, int p_grid, int q_grid, int N,
                          int nZobs, int nZmiss);
/* exageostat_exact/src/include/MLE_exact_s.h:59-61, MLE_sdexact.h:63-65,
 * exageostat_approx/src/include/MLE_approx.h:54-56: the same device context
 * as EXAGEOSTAT_dmle_Call (the precision of each tile is chosen per
 * evaluation). */
void EXAGEOSTAT_smle_Call(MLE_data *data, int ncores, int gpus, int dts, int p_grid, int q_grid, int N, int nZobs,
                          int nZmiss);
void EXAGEOSTAT_sdmle_Call(MLE_data *data, int ncores, int gpus, int dts, int p_grid, int q_grid, int N, int nZobs,
                           int nZmiss);
void EXAGEOSTAT_dmle_diag_Call(MLE_data *data, int ncores, int gpus, int dts, int p_grid, int q_grid, int N,
                               int nZobs, int nZmiss);
/* src/compute/MLE.c:381-401. */
void MLE_Finalize(MLE_data *data);
/* src/compute/MLE.c:420-434. */
void MLE_get_zobs(MLE_data *data, double *z, int n);

/* ------------------------------------------------------ synthetic data -- */
/* misc/compute/MLE_misc.c:763-792 (srand(seed), jittered grid, Morton sort). */
location *GenerateXYLoc(int n, int seed);
/* misc/compute/MLE_misc.c:887-923: n jittered-grid points (srand(seed), x
 * drawn before y, no Morton sort) repeated for t_slots time slots; location
 * i + t n has z = t + 1 (the time of the space-time kernel). */
location *GenerateXYLoc_ST(int n, int t_slots, int seed);
/* misc/compute/MLE_misc.c:1700-1741. */
void pick_random_points(MLE_data *data, double *Zobs, double *Zactual, int nZmiss, int nZobs, int N);
/* src/compute/MLE.c:38-92 -> EXAGEOSTAT_MLE_dzvg_Tile, MLE_dexact.c:25-148:
 * z = L(theta) e on the GPU. */
void MLE_zvg(MLE_data *data, double *Nrand, double *initial_theta, int n, int ts, int log, int p_grid,
             int q_grid);
void EXAGEOSTAT_MLE_dzvg_Tile(MLE_data *data, double *Nrand, double *initial_theta, int n, int dts, int log);
/* MLE_exact.h:38-39 (MLE_dexact.c:167-229); stream-ordered, same result. */
void EXAGEOSTAT_MLE_dzvg_Tile_Async(MLE_data *data, double *Nrand, double *initial_theta, int n, int dts, int log);

/* ------------------------------------------------------------ likelihood -- */
/* src/include/MLE.h:48-49, src/compute/MLE.c:138-193: NLOPT objective.
 * Dispatches as the reference does: computation "exact" with precision 0
 * (EXAGEOSTAT_dmle_Tile), 1 (EXAGEOSTAT_smle_Tile) or 2
 * (EXAGEOSTAT_sdmle_Tile); "diag_approx" with precision 0
 * (EXAGEOSTAT_dmle_diag_Tile).  Other combinations return -HUGE_VAL with
 * status -1 (the reference returns 0).  grad, when non-NULL, receives
 * d l / d theta on the precision-0 exact path and on the precision-2
 * mixed path (the reference ignores it; the mixed gradient takes alpha
 * from the refined solve and the trace term from the mixed factor, and
 * needs a context that holds the FP64 covariance); precision 1 and
 * "diag_approx" then fail with status -1.  The "_profile" bivariate
 * kernels return the profile likelihood and set variance1 / variance2
 * (MLE_dexact.c:403-421).  With a non-empty recovery_file, an evaluation
 * whose iteration number has a line there returns that line's likelihood
 * without computing (MLE_dexact.c:306-307); with a non-empty
 * checkpoint_file every computed evaluation appends its line. */
double MLE_alg(unsigned n, const double *theta, double *grad, void *data);
/* exageostat_exact/src/include/MLE_exact.h:43-47, MLE_dexact.c:231-611.
 * The _Async variant is stream-ordered here and returns the same value. */
double EXAGEOSTAT_dmle_Tile(unsigned n, const double *theta, double *grad, void *data);
double EXAGEOSTAT_dmle_Tile_Async(unsigned n, const double *theta, double *grad, void *data);
double MORSE_dmle_Tile(unsigned n, const double *theta, double *grad, void *data);
/* exageostat_exact/src/include/MLE_sdexact.h:45-49 (EXAGEOSTAT_sdmle_Tile,
 * MLE_sdexact.c:172-357): mixed precision.  Here: FP64 diagonal tiles, FP32 /
 * BF16 off-diagonal tiles chosen per theta, FP64 refinement of the solve. */
double EXAGEOSTAT_sdmle_Tile(unsigned n, const double *theta, double *grad, void *data);
double EXAGEOSTAT_sdmle_Tile_Async(unsigned n, const double *theta, double *grad, void *data);
/* exageostat_exact/src/include/MLE_exact_s.h:41-45 (EXAGEOSTAT_smle_Tile,
 * MLE_sexact.c:155-277): single precision.  Every off-diagonal tile is
 * stored in FP32 and updated with FP32-accurate products (split BF16 on the
 * matrix cores); the diagonal tiles, their factorisation and the solve
 * arithmetic are FP64 (the diagonal-tile kernels are FP64 only); no
 * refinement.  data->sdotp receives the dot product in FP32. */
double EXAGEOSTAT_smle_Tile(unsigned n, const double *theta, double *grad, void *data);
double EXAGEOSTAT_smle_Tile_Async(unsigned n, const double *theta, double *grad, void *data);
/* exageostat_approx/src/include/MLE_approx.h:41-45 (EXAGEOSTAT_dmle_diag_Tile,
 * MLE_approx.c:201-365): the band-limited factorisation of
 * CHAM_pdpotrf_diag (pdpotrf_diag.c:40-110) with band data->diag_thick
 * tiles, then the log-determinant over the diagonal tiles and the forward
 * solve over the whole lower triangle (whose tiles outside the band still
 * hold the covariance), as there. */
double EXAGEOSTAT_dmle_diag_Tile(unsigned n, const double *theta, double *grad, void *data);
double EXAGEOSTAT_dmle_diag_Tile_Async(unsigned n, const double *theta, double *grad, void *data);

/* ------------------------------------------------------------ tile tasks -- */
/* runtime/starpu/include/starpu_exageostat.h:52-55, codelet_zcmg.c:246-347. */
int EXAGEOSTAT_MLE_dcmg_Tile_Async(CHAM_enum uplo, CHAM_desc_t *descA, location *l1, location *l2,
                                   location *lm, double *theta, char *dm, char *kernel_fun,
                                   RUNTIME_sequence_t *sequence, RUNTIME_request_t *request);
int MORSE_MLE_dcmg_Tile_Async(CHAM_enum uplo, CHAM_desc_t *descA, location *l1, location *l2, location *lm,
                              double *theta, char *dm, char *kernel_fun, RUNTIME_sequence_t *sequence,
                              RUNTIME_request_t *request);
/* starpu_exageostat.h:74-76, codelet_zmdet.c:108-139: *descdet->mat +=
 * sum log L_ii over positive pivots (after EXAGEOSTAT_Sequence_Wait). */
int EXAGEOSTAT_MLE_dmdet_Tile_Async(CHAM_desc_t *descA, RUNTIME_sequence_t *sequence,
                                    RUNTIME_request_t *request, CHAM_desc_t *descdet);
int MORSE_MLE_dmdet_Tile_Async(CHAM_desc_t *descA, RUNTIME_sequence_t *sequence, RUNTIME_request_t *request,
                               CHAM_desc_t *descdet);

/* ------------------------------------------------------------ prediction -- */
/* MLE_exact.h:49-55, MLE_dexact.c:613-685 / 688-874. */
void EXAGEOSTAT_dmle_Predict_Allocate(MLE_data *data, int nZmiss, int nZobs, int dts, int p_grid, int q_grid,
                                      int mse_flag);
double EXAGEOSTAT_dmle_Predict_Tile(MLE_data *data, double *theta, int nZmiss, int nZobs, double *Zobs,
                                    double *Zactual, double *Zmiss, int n);
double MORSE_dmle_Predict_Tile(MLE_data *data, double *theta, int nZmiss, int nZobs, double *Zobs,
                               double *Zactual, double *Zmiss, int n);
/* MLE_exact.h:55-56 (MLE_dexact.c:876-1000): held-out locations = the first
 * nZmiss of data->l1, observed = the rest; Zobs and Zactual from
 * data->descZobs / data->descZactual (set by Predict_Allocate, filled by
 * the caller with EXAGEOSTAT_Lapack_to_Tile), predictions into
 * data->descZmiss; returns the MSE (-1 without actual values). */
double EXAGEOSTAT_dmle_Predict_Tile_Async(MLE_data *data, double *theta, int nZmiss, int nZobs, int n);
/* MLE_exact_s.h:47-50 (single precision: C22 factored with FP32 off-diagonal
 * tiles, solved with that factor), MLE_sdexact.h:51-54 (mixed precision: C22
 * factored under the mp_* options, the solve refined in FP64; status -3
 * when the refinement does not reach mp_ir_tol, the predictions still
 * returned), MLE_approx.h:47-50 (the FP64 predictor).  C12 and the product
 * stay FP64 in all three. */
double EXAGEOSTAT_smle_Predict_Tile(MLE_data *data, double *theta, int nZmiss, int nZobs, double *Zobs,
                                    double *Zactual, double *Zmiss, int n);
double EXAGEOSTAT_sdmle_Predict_Tile(MLE_data *data, double *theta, int nZmiss, int nZobs, double *Zobs,
                                     double *Zactual, double *Zmiss, int n);
double EXAGEOSTAT_dmle_diag_Predict_Tile(MLE_data *data, double *theta, int nZmiss, int nZobs, double *Zobs,
                                         double *Zactual, double *Zmiss, int n);
/* src/compute/MLE.c:284-306, 341-380. */
void prediction_init(MLE_data *data, int nZmiss, int nZobs, int ts, int p_grid, int q_grid, int mse_flag);
void prediction_finalize(MLE_data *data);

/* ---------------------------------------------------------- MLOE / MMOM -- */
/* MLE_exact.h:69-76, MLE_dexact.c:1002-1435; src/compute/MLE.c:309-320,403.
 * Prediction efficiency of kriging with estimated_theta when truth_theta
 * holds, over the held-out locations data->lmiss (nZmiss) given
 * data->lobs (nZobs); univariate Matern kernels.  Results in data->mloe and
 * data->mmom (the means over the held-out locations of E2/E3 - 1 and
 * E4/E2 - 1, E2: true MSE of the estimated predictor, E3: true MSE of the
 * optimal predictor, E4: MSE presumed under estimated_theta); data->status
 * is nonzero when a covariance matrix is not positive definite.  The
 * _Async variant is synchronous here. */
void EXAGEOSTAT_dmle_mloe_mmom_Allocate(MLE_data *data, int nZmiss, int nZobs, int dts, int p_grid, int q_grid);
void EXAGEOSTAT_dmle_mloe_mmom_Tile(MLE_data *data, double *truth_theta, double *estimated_theta, int nZmiss,
                                    int nZobs, int n);
void EXAGEOSTAT_dmle_mloe_mmom_Tile_Async(MLE_data *data, double *truth_theta, double *estimated_theta, int nZmiss,
                                          int nZobs, int n);
void mloe_mmom_init(MLE_data *data, int nZmiss, int nZobs, int ts, int p_grid, int q_grid);
void mloe_mmom_finalize(MLE_data *data);
void MLOE_MMOM_Finalize(MLE_data *data);
/* No reference counterpart: the same computation returning the per-location
 * values loe[j] = E2/E3 - 1 and mom[j] = E4/E2 - 1 (each array nZmiss, may be
 * NULL).  Returns 0, > 0 when a covariance matrix is not positive definite,
 * -1 on an invalid argument. */
int exageostat_mloe_mmom(MLE_data *data, const double *truth_theta, const double *estimated_theta, int nZmiss,
                         int nZobs, double *loe, double *mom);

/* ------------------------------------------------- Fisher information -- */
/* exageostat_exact/src/compute/MLE_dexact.c:1777-1934 (declared in
 * MLE_exact.h).  Fisher information F_jk = 1/2 tr(Sigma^-1 dSigma_j
 * Sigma^-1 dSigma_k) at initial_theta over data->l1 (N locations; dts,
 * p_grid, q_grid as in the reference, the context's tile size is used);
 * returns a malloc'd array of np*np + 1 doubles holding F^-1 column-major
 * (np = 3 for "univariate_matern_stationary" as in the reference, 4 with
 * the nugget kernel), or NULL (status in exageostat_gpu_last_status). */
double *EXAGEOSTAT_Fisher_Tile(MLE_data *data, int N, double *initial_theta, int dts, int p_grid, int q_grid);
/* No counterpart in the reference.  F and F^-1 (np x np each, symmetric)
 * at theta; returns 0, > 0 (Sigma(theta) not positive definite: first
 * failing column + 1), -1 (error) or -2 (F not positive definite; F^-1 is
 * then NaN). */
int exageostat_fisher(MLE_data *data, const double *theta, double *fisher, double *fisher_inv);

/* ------------------------------------------- descriptors and sequences -- */
/* Stand-ins for the Chameleon calls the reference makes on this path
 * (CHAMELEON_Desc_Create via EXAGEOSTAT_ALLOCATE_MATRIX_TILE, MLE_misc.h:81-87;
 * CHAMELEON_Sequence_Create/Wait; CHAMELEON_Lapack_to_Tile/Tile_to_Lapack;
 * CHAMELEON_dpotrf_Tile, CHAMELEON_dtrsm_Tile). */
int EXAGEOSTAT_Desc_Create(CHAM_desc_t **desc, CHAM_enum uplo, int mb, int m, int n);
/* A 1x1 (or m x n) descriptor over caller host memory, as the reference's
 * descdet / descproduct alias &data->det / &data->dotp (MLE_dexact.c:1981-1988). */
int EXAGEOSTAT_Desc_Create_User(CHAM_desc_t **desc, double *host_mat, int m, int n);
int EXAGEOSTAT_Desc_Destroy(CHAM_desc_t **desc);
int EXAGEOSTAT_Sequence_Create(RUNTIME_sequence_t **sequence);
int EXAGEOSTAT_Sequence_Wait(RUNTIME_sequence_t *sequence);
int EXAGEOSTAT_Sequence_Destroy(RUNTIME_sequence_t *sequence);
int EXAGEOSTAT_Lapack_to_Tile(const double *A, int lda, CHAM_desc_t *desc);
int EXAGEOSTAT_Tile_to_Lapack(CHAM_desc_t *desc, double *A, int lda);
/* In-place lower Cholesky; returns 0 or the first failing column (1-based). */
int EXAGEOSTAT_dpotrf_Tile(CHAM_enum uplo, CHAM_desc_t *A);
/* Sigma^-1 in place of the factor from EXAGEOSTAT_dpotrf_Tile (the last
 * matrix factored): lower tiles, diagonal tiles written in full.  Replaces
 * CHAMELEON_dpotri_Tile (a dependency of the reference, not vendored:
 * chameleon submodule, .gitmodules:4-6), used here by the gradient. */
int EXAGEOSTAT_dpotri_Tile(CHAM_enum uplo, CHAM_desc_t *A);
/* B <- L^-1 B (trans = ChamNoTrans) or L^-T B (ChamTrans), B an n-vector
 * given in host memory; L from EXAGEOSTAT_dpotrf_Tile. */
int EXAGEOSTAT_dtrsv_Tile(CHAM_enum trans, CHAM_desc_t *L, double *b);
/* No counterpart in the reference.  EXAGEOSTAT_dtrsv_Tile without any
 * device allocation, for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * its scratch (the padded right-hand side, the solve's assembly vector,
 * partial sums, a completion counter) is a static device array of 1.13 MB
 * per GPU, held until the process ends.  Limits: mt * ts <= 65536 and
 * ts <= 1024; larger L returns -1 with an error string.  Every call on one
 * GPU shares that array: calls that overlap on one GPU (other host threads)
 * race.  L must come from EXAGEOSTAT_dpotrf_Tile (its inverted diagonal
 * tiles are read).  Returns 0 or -1. */
int EXAGEOSTAT_dtrsv_Tile_noalloc(CHAM_enum trans, CHAM_desc_t *L, double *b);

/* LAPACK dlarnv (the reference draws N(0,1) with LAPACKE_dlarnv(3, iseed, N,
 * Nrand), examples/synthetic_dmle_test.c:111,147); iseed is advanced. */
void exageostat_dlarnv(int idist, int *iseed, int n, double *x);

/* ------------------------------------------------------------- optimiser -- */
/* The objective type the reference registers with NLOPT (nlopt_func;
 * MLE_alg has this signature). */
typedef double exageostat_func_t(unsigned n, const double *x, double *grad, void *data);
/* Replaces the reference's NLOPT_LN_BOBYQA run (examples/
 * synthetic_dmle_test.c:140-142,283-284; init_optimizer, misc/compute/
 * MLE_misc.c:1352-1366), which needs libnlopt: maximises f over the box
 * [lb, ub] from x (overwritten with the best point) by Powell's BOBYQA
 * (2009) with 2n+1 interpolation points, NLOPT's default initial step and
 * variable scaling, variables with lb == ub held fixed; stops when an
 * evaluation after the initial ones lies within ftol_abs of the best value
 * (nlopt_set_ftol_abs), after maxeval evaluations (maxeval <= 0: no limit,
 * as NLOPT), or when the trust region has shrunk to 1e-14 of the initial
 * step.  *fmax: best value, *nevals: evaluations.  Returns 0 (ftol), 1
 * (maxeval), 2 (step floor) or < 0 (invalid arguments, degenerate points). */
int exageostat_bobyqa(unsigned n, exageostat_func_t *f, void *f_data, const double *lb, const double *ub, double *x,
                      double ftol_abs, int maxeval, double *fmax, int *nevals);

/* ------------------------------------------------- real data and R .C -- */
/* misc/compute/MLE_misc.c:24-144: MLE_data fields to their defaults (this
 * implementation's fields included). */
void init_data_values(MLE_data *data);
/* misc/compute/flat_file.c:23-63, 65-111: "x,y" (or "x,y,z") per line, n
 * locations; NULL if the file cannot be opened (the 3-D reader of the
 * reference exits).  The caller frees x, y (z) and the struct. */
location *readLocsFile(char *locs_file, int n);
location *readLocsFile3d(char *locs_file, int n);
/* flat_file.c:115-145, 147-175: one value per line; malloc'd array of n
 * doubles, or NULL if the file cannot be opened (the reference exits). */
double *readObsFile(char *obsfile, int n);
double *readTimeFile(char *timefile, int n);
/* The R package's .C entry points (r-wrappers/compute/rwrappers.c), every
 * argument a pointer.  Kernels: 0 univariate Matern, 1 with nugget, 3
 * parsimonious bivariate, 5 space-time (generation only; 2, 4 and 6 are
 * not supported).  dmetric 0 Euclidean, 1 great circle.  The maximum-
 * likelihood calls run exageostat_bobyqa over MLE_alg from clb with
 * ftol_abs = *opt_tol and at most *opt_max_iters evaluations (<= 0: no
 * limit) and write [theta hat, seconds per evaluation, total seconds,
 * evaluations] to globalthetaout (theta hat -1 on failure).  :1172-1187. */
void rexageostat_init(int *ncores, int *gpus, int *dts, int *lts);
void rexageostat_finalize(void);
/* :56-176: z = L(theta) e at the given locations, e = dlarnv(3, {0,0,0,1});
 * globalvec (veclen >= rows) receives z. */
void gen_z_givenlocs_exact(double *x, int *xlen, double *y, int *ylen, int *kernel, double *theta, int *thetalen,
                           int *dmetric, int *n, int *ncores, int *gpus, int *ts, int *p_grid, int *q_grid,
                           int *veclen, double *globalvec);
/* :178-301: GenerateXYLoc(n, seed), e = dlarnv(3, {seed,seed,seed,1});
 * globalvec = [x (n), y (n), z (rows)]. */
void gen_z_exact(int *kernel, double *theta, int *thetalen, int *dmetric, int *n, int *seed, int *ncores, int *gpus,
                 int *ts, int *p_grid, int *q_grid, int *veclen, double *globalvec);
/* :710-753 (computation "exact", precision 0). */
void mle_exact(double *x, int *xlen, double *y, int *ylen, double *z, int *zlen, double *clb, int *clblen, double *cub,
               int *cublen, int *kernel, int *dmetric, int *n, double *opt_tol, int *opt_max_iters, int *ncores,
               int *gpus, int *ts, int *p_grid, int *q_grid, double *globalthetaout);
/* :848-891 ("diag_approx" with band dst_thick). */
void mle_dst(double *x, int *xlen, double *y, int *ylen, double *z, int *zlen, int *kernel, double *clb, int *clblen,
             double *cub, int *cublen, int *dst_thick, int *dmetric, int *n, double *opt_tol, int *opt_max_iters,
             int *ncores, int *gpus, int *ts, int *p_grid, int *q_grid, double *globalthetaout);
/* :893-936 (mixed precision, FP64 band mp_band; the reference evaluates
 * these in FP64 because its mle_general leaves precision at 0). */
void mle_mp(double *x, int *xlen, double *y, int *ylen, double *z, int *zlen, int *kernel, double *clb, int *clblen,
            double *cub, int *cublen, int *mp_band, int *dmetric, int *n, double *opt_tol, int *opt_max_iters,
            int *ncores, int *gpus, int *ts, int *p_grid, int *q_grid, double *globalthetaout);
/* :938-1072: kriging; globalzpredict receives nZmiss (x2 for bivariate)
 * predictions (-1000 where nothing was computed); computation must be 0. */
void exact_predict(double *xobs, int *xobs_len, double *yobs, int *yobs_len, double *zobs, int *zobs_len,
                   double *xmiss, int *xmiss_len, double *ymiss, int *ymiss_len, int *nZobs, int *nZmiss, int *kernel,
                   double *est_theta, int *thetalen, int *computation, int *dmetric, int *ncores, int *gpus, int *ts,
                   int *p_grid, int *q_grid, double *globalzpredict);
/* :1074-1170: global_mloe_mmom = [MLOE, MMOM] (-1000 on failure). */
void exact_mloe_mmom(double *xobs, int *xobs_len, double *yobs, int *yobs_len, double *zobs, int *zobs_len,
                     double *xmiss, int *xmiss_len, double *ymiss, int *ymiss_len, int *nZobs, int *nZmiss,
                     int *kernel, double *est_theta, double *true_theta, int *thetalen, int *dmetric, int *ncores,
                     int *gpus, int *ts, int *p_grid, int *q_grid, double *global_mloe_mmom);

/* ------------------------------------------------------------ measurement -- */
/* No counterpart in the reference.  Per-kernel HIP-event timing on the
 * launching stream ("tile_update", "potrf_diag", "trsm_panel", "gen", "trsv"),
 * used by the benchmark; off by default. */
void exageostat_gpu_synchronize(void);
/* No counterpart in the reference.  Tile precision classes (0 FP64, 1 FP32,
 * 3 FP32 storage with BF16 updates) of the last mixed-precision evaluation
 * (precision = 2), lower-packed column by column; returns the tile count. */
long exageostat_mixed_prec_map(MLE_data *data, unsigned char *out, long cap);
void exageostat_gpu_profile_enable(int on);
void exageostat_gpu_profile_reset(void);
int exageostat_gpu_profile_get(const char *name, long *launches, double *ms, double *flops, double *bytes);

/* ------------------------------------------------------------- status -- */
int exageostat_gpu_last_status(void);
const char *exageostat_gpu_last_error(void);
/* Number of HIP devices visible (0 without a GPU). */
int exageostat_gpu_device_count(void);

#ifdef __cplusplus
}
#endif

#endif /* EXAGEOSTAT_GPU_H */
