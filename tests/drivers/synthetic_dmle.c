/*
 * A C caller of libexageostat_gpu.so written the way the reference's
 * examples/synthetic_dmle_test.c drives its own library: the MLE_data
 * fields its argument parsing sets (examples/examples.c), GenerateXYLoc,
 * dlarnv, exageostat_init, EXAGEOSTAT_dmle_Call, MLE_zvg, MLE_alg (here over
 * a list of theta read from stdin in place of the NLOPT loop, which needs
 * libnlopt), pick_random_points, prediction_init and
 * EXAGEOSTAT_dmle_Predict_Tile, then MLE_Finalize.
 *
 * Usage: synthetic_dmle N TS SEED NZMISS KERNEL < thetas
 * (one line of num_params values per evaluation).  Prints one line
 * "loglik <l> <logdet> <dotp> <status>" per theta, then "zmiss <values>",
 * "mse <value>" and "results <final_loglik>".  Exit status 0 on success.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "exageostat_gpu.h"

static int num_params_of(const char *kernel) {
  if (!strcmp(kernel, "univariate_matern_stationary")) return 3;
  if (!strcmp(kernel, "univariate_matern_nuggets_stationary")) return 4;
  return -1;
}

int main(int argc, char **argv) {
  if (argc < 6) {
    fprintf(stderr, "usage: %s N TS SEED NZMISS KERNEL < thetas\n", argv[0]);
    return 2;
  }
  int N = atoi(argv[1]);
  int dts = atoi(argv[2]);
  int seed = atoi(argv[3]);
  int nZmiss = atoi(argv[4]);
  char *kernel = argv[5];
  int np = num_params_of(kernel);
  if (np < 0) {
    fprintf(stderr, "unsupported kernel %s\n", kernel);
    return 2;
  }
  int nZobs = N - nZmiss;

  /* the fields examples.c sets from its arguments */
  MLE_data data;
  memset(&data, 0, sizeof(data));
  data.computation = "exact";
  data.kernel_fun = kernel;
  data.dm = "ed";
  data.c_fun = "pow-exp";
  data.precision = 0;
  data.async = 0;
  data.check = 0;
  data.verbose = 0;
  data.log = 0;
  data.ooc = 0;
  data.fisher = 0;
  data.idw = 0;
  data.mspe = 1;
  data.mloe_mmom = 0;
  data.mloe_mmom_async = 0;
  data.time_slots = 0;
  data.diag_thick = 1;
  data.opt_tol = 5;
  data.opt_max_iters = 100;
  data.recovery_file = "";
  data.checkpoint_file = "";
  data.actualZFPath = "";
  data.seed = seed;
  data.silent = 1;

  int ncores = 1, gpus = 1, lts = 0, p_grid = 1, q_grid = 1;
  double *initial_theta = malloc(sizeof(double) * np);
  initial_theta[0] = 1.0;
  initial_theta[1] = 0.1;
  initial_theta[2] = 0.5;
  if (np == 4) initial_theta[3] = 0.01;

  results.problem_size = N;
  exageostat_init(&ncores, &gpus, &dts, &lts);
  if (exageostat_gpu_last_status() != 0) {
    fprintf(stderr, "exageostat_init: %s\n", exageostat_gpu_last_error());
    return 1;
  }
  int iseed[4] = {seed, seed, seed, 1};
  double *Nrand = malloc(sizeof(double) * N);
  exageostat_dlarnv(3, iseed, N, Nrand);
  location *locations = GenerateXYLoc(N, seed);
  data.l1 = *locations;

  EXAGEOSTAT_dmle_Call(&data, ncores, gpus, dts, p_grid, q_grid, N, nZobs, nZmiss);
  if (exageostat_gpu_last_status() != 0) {
    fprintf(stderr, "EXAGEOSTAT_dmle_Call: %s\n", exageostat_gpu_last_error());
    return 1;
  }
  MLE_zvg(&data, Nrand, initial_theta, N, dts, 0, p_grid, q_grid);
  if (exageostat_gpu_last_status() != 0) {
    fprintf(stderr, "MLE_zvg: %s\n", exageostat_gpu_last_error());
    return 1;
  }

  double *theta = malloc(sizeof(double) * np);
  results.estimated_theta = malloc(sizeof(double) * np);
  for (;;) {
    int got = 0;
    for (int i = 0; i < np; ++i) got += scanf("%lf", &theta[i]);
    if (got != np) break;
    double l = MLE_alg((unsigned)np, theta, NULL, &data);
    printf("loglik %.17g %.17g %.17g %d\n", l, 2.0 * data.det, data.dotp, data.status);
  }

  if (nZmiss > 0) {
    double *Zobs = malloc(sizeof(double) * nZobs);
    double *Zactual = malloc(sizeof(double) * nZmiss);
    double *Zmiss = malloc(sizeof(double) * nZmiss);
    pick_random_points(&data, Zobs, Zactual, nZmiss, nZobs, N);
    prediction_init(&data, nZmiss, nZobs, dts, p_grid, q_grid, 1);
    double mse = EXAGEOSTAT_dmle_Predict_Tile(&data, initial_theta, nZmiss, nZobs, Zobs, Zactual, Zmiss, N);
    printf("zmiss");
    for (int i = 0; i < nZmiss; ++i) printf(" %.17g", Zmiss[i]);
    printf("\nmse %.17g %.17g\n", mse, results.mse_pred);
    prediction_finalize(&data);
    free(Zobs);
    free(Zactual);
    free(Zmiss);
  }
  printf("results %.17g %d\n", results.final_loglik, data.iter_count);
  MLE_Finalize(&data);
  exageostat_finalize();
  free(theta);
  free(Nrand);
  free(initial_theta);
  return exageostat_gpu_last_status() == 0 ? 0 : 1;
}
