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
 *  - MLE_data is the subset of the reference struct
 *    (misc/include/MLE_misc.h:184-328) that the path reads or writes, plus
 *    a few fields of this implementation at the end.
 */
#ifndef EXAGEOSTAT_GPU_H
#define EXAGEOSTAT_GPU_H

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

/* Subset of MLE_data (misc/include/MLE_misc.h:184-328); field names and
 * meanings as in the reference. */
typedef struct {
  double variance;
  double variance1;
  double variance2;
  char *computation;    /* "exact"                                          */
  int test;
  int async;
  int iter_count;
  location l1;          /* locations of the observations                    */
  location lmiss;       /* prediction: held-out locations                   */
  location lobs;        /* prediction: observed locations                   */
  location lm;
  void *descC;          /* Sigma, lower tiles                               */
  void *descZ;
  void *descZcpy;
  void *descdet;
  void *descproduct;
  void *descZmiss;
  void *descC12;
  void *descC22;
  void *descZactual;
  void *descZobs;
  void *descmse;
  void *sequence;
  void *request;
  int verbose;
  int log;
  double avg_exec_time_per_iter;
  double total_exec_time;
  double avg_flops_per_iter;
  double avg_generation_per_iter;
  double avg_solve_per_iter;
  double avg_cholesky_per_iter;
  double final_loglik;
  char *recovery_file;
  char *checkpoint_file;
  double det;           /* sum of log L_ii (logdet = 2*det)                  */
  double dotp;          /* ||L^-1 z||^2                                      */
  double mserror;
  double mserror1;
  double mserror2;
  char *dm;             /* "ed" Euclidean, "gc" great circle                 */
  int diag_thick;       /* mixed precision: tile band kept in FP64           */
  int opt_tol;
  int opt_max_iters;
  char *kernel_fun;     /* "univariate_matern_stationary", ...              */
  int precision;        /* 0 double, 2 mixed (FP64 / FP32 / BF16 tiles)      */
  int mspe;
  int seed;
  double mloe;          /* MLOE of the last EXAGEOSTAT_dmle_mloe_mmom_Tile     */
  double mmom;          /* MMOM of the last EXAGEOSTAT_dmle_mloe_mmom_Tile     */
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
  int mp_ir_steps;      /* mixed: FP64 iterative-refinement steps on the solve */
  double mp_fp64_corr;  /* mixed: off-diagonal tiles whose largest
                           correlation is >= this stay FP64                  */
  int mp_lolo;          /* mixed: 1 adds the lo*lo term to the split-BF16
                           products of the FP32 tiles (four products instead
                           of three: smaller log-det bias, ~6% slower)       */
} MLE_data;

/* ---------------------------------------------------------------- setup -- */
/* src/compute/MLE.c:233-260 (CHAMELEON_Init(ncores, gpus)). */
void exageostat_init(int *ncores, int *gpus, int *dts, int *lts);
/* src/compute/MLE.c (CHAMELEON_Finalize). */
void exageostat_finalize(void);
/* exageostat_exact/src/compute/MLE_dexact.c:1938-2017. */
void EXAGEOSTAT_dmle_Call(MLE_data *data, int ncores, int gpus, int dts, int p_grid, int q_grid, int N,
                          int nZobs, int nZmiss);
/* src/compute/MLE.c:381-401. */
void MLE_Finalize(MLE_data *data);
/* src/compute/MLE.c:420-434. */
void MLE_get_zobs(MLE_data *data, double *z, int n);

/* ------------------------------------------------------ synthetic data -- */
/* misc/compute/MLE_misc.c:763-792 (srand(seed), jittered grid, Morton sort). */
location *GenerateXYLoc(int n, int seed);
/* misc/compute/MLE_misc.c:1700-1741. */
void pick_random_points(MLE_data *data, double *Zobs, double *Zactual, int nZmiss, int nZobs, int N);
/* src/compute/MLE.c:38-92 -> EXAGEOSTAT_MLE_dzvg_Tile, MLE_dexact.c:25-148:
 * z = L(theta) e on the GPU. */
void MLE_zvg(MLE_data *data, double *Nrand, double *initial_theta, int n, int ts, int log, int p_grid,
             int q_grid);
void EXAGEOSTAT_MLE_dzvg_Tile(MLE_data *data, double *Nrand, double *initial_theta, int n, int dts, int log);

/* ------------------------------------------------------------ likelihood -- */
/* src/include/MLE.h:48-49, src/compute/MLE.c:138-193: NLOPT objective.
 * grad, when non-NULL, receives d l / d theta (the reference ignores it). */
double MLE_alg(unsigned n, const double *theta, double *grad, void *data);
/* exageostat_exact/src/include/MLE_exact.h:43-44, MLE_dexact.c:231-489. */
double EXAGEOSTAT_dmle_Tile(unsigned n, const double *theta, double *grad, void *data);
double MORSE_dmle_Tile(unsigned n, const double *theta, double *grad, void *data);
/* exageostat_exact/src/include/MLE_sdexact.h (EXAGEOSTAT_sdmle_Tile,
 * MLE_sdexact.c:172-357): mixed precision.  Here: FP64 diagonal tiles, FP32 /
 * BF16 off-diagonal tiles chosen per theta, FP64 refinement of the solve. */
double EXAGEOSTAT_sdmle_Tile(unsigned n, const double *theta, double *grad, void *data);

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

/* LAPACK dlarnv (the reference draws N(0,1) with LAPACKE_dlarnv(3, iseed, N,
 * Nrand), examples/synthetic_dmle_test.c:111,147); iseed is advanced. */
void exageostat_dlarnv(int idist, int *iseed, int n, double *x);

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
