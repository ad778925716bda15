// The reference's second caller of MLE_alg: the R `.C` entry points
// (r-wrappers/compute/rwrappers.c) and the flat-file readers of real data
// (misc/compute/flat_file.c), written over this library's C ABI.  Every
// argument is a pointer, as R's .C interface passes them.  The NLOPT BOBYQA
// run of the reference's mle_general is exageostat_bobyqa here, with
// MLE_alg as the objective and the MLE_data pointer as its data, the same
// bounds, start (the lower bounds), ftol_abs (*opt_tol) and maximum number of
// evaluations.
//
// Deliberate differences, all where the reference does not do what it
// says:
//  - dmetric 1 selects the great-circle distance ("gc"); the reference
//    writes "gcd", which its generator does not recognise, so its R calls
//    always use Euclidean distances (rwrappers.c:608 against
//    codelet_zcmg.c:260);
//  - mle_mp runs the mixed-precision path (precision 2) with the FP64 band
//    mp_band; the reference's mle_general leaves precision at 0 and so
//    evaluates in FP64 (rwrappers.c:640);
//  - exact_predict and exact_mloe_mmom size theta from *thetalen (the
//    reference allocates num_params = 0 doubles and then writes thetalen of
//    them, rwrappers.c:993-1000);
//  - errors return with the outputs left at their initial markers (-1 or
//    -1000) instead of exit().
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/exageostat_gpu.h"

namespace {

const char* kernel_name(int k) {
  switch (k) {
    case 0: return "univariate_matern_stationary";
    case 1: return "univariate_matern_nuggets_stationary";
    case 3: return "bivariate_matern_parsimonious";
    case 5: return "univariate_spacetime_matern_stationary";
    default: return nullptr;  // 2, 4, 6: flexible bivariate, trivariate, bivariate space-time
  }
}

int kernel_params(int k) {
  switch (k) {
    case 0: return 3;
    case 1: return 4;
    case 3: return 6;
    case 5: return 7;
    default: return -1;
  }
}

int kernel_vars(int k) { return k == 3 ? 2 : 1; }

bool check_kernel(int k, const char* fn) {
  if (kernel_name(k)) return true;
  fprintf(stderr, "%s: kernel %d is not supported by this library\n", fn, k);
  return false;
}

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Reads the whole file line by line; returns false if it cannot be opened.
bool read_lines(const char* path, std::vector<std::string>& out) {
  FILE* fp = fopen(path, "r");
  if (!fp) return false;
  char* line = nullptr;
  size_t cap = 0;
  while (getline(&line, &cap, fp) != -1) out.emplace_back(line);
  free(line);
  fclose(fp);
  return true;
}

}  // namespace

extern "C" {

void init_data_values(MLE_data* data) {
  // misc/compute/MLE_misc.c:24-144: every field to its default; this
  // implementation's own fields as a fresh MLE_data needs them.
  memset(data, 0, sizeof(*data));
  data->variance = data->variance1 = data->variance2 = data->variance3 = 1.0;
  data->computation = (char*)"exact";
  data->c_fun = (char*)"univariate_matern_stationary";
  data->dm = (char*)"ed";
  data->kernel_fun = (char*)"univariate_matern_stationary";
  data->diag_thick = 1;
  data->opt_tol = 5;
  data->opt_max_iters = -1;
  data->recovery_file = (char*)"";
  data->checkpoint_file = (char*)"";
  data->obsFPath = (char*)"";
  data->locsFPath = (char*)"";
  data->timeFPath = (char*)"";
  data->actualZFPath = (char*)"";
  data->silent = 1;
  data->mp_ir_steps = 3;
}

// -------------------------------------------------------------- flat files
// misc/compute/flat_file.c:23-63: one "x,y" line per location.
location* readLocsFile(char* locs_file, int n) {
  std::vector<std::string> lines;
  if (!read_lines(locs_file, lines)) {
    printf("cannot read locations file\n%s: \n", locs_file);
    return nullptr;
  }
  auto* loc = static_cast<location*>(malloc(sizeof(location)));
  loc->x = static_cast<double*>(calloc(n > 0 ? n : 1, sizeof(double)));
  loc->y = static_cast<double*>(calloc(n > 0 ? n : 1, sizeof(double)));
  loc->z = nullptr;
  for (int i = 0; i < n && i < (int)lines.size(); ++i) {
    char* save = nullptr;
    std::vector<char> buf(lines[i].begin(), lines[i].end());
    buf.push_back('\0');
    const char* a = strtok_r(buf.data(), ",", &save);
    const char* b = a ? strtok_r(nullptr, ",", &save) : nullptr;
    loc->x[i] = a ? atof(a) : 0.0;
    loc->y[i] = b ? atof(b) : 0.0;
  }
  return loc;
}

// flat_file.c:65-111: "x,y,z" per line.
location* readLocsFile3d(char* locs_file, int n) {
  std::vector<std::string> lines;
  if (!read_lines(locs_file, lines)) {
    printf("cannot open location file%s\n", locs_file);
    return nullptr;
  }
  auto* loc = static_cast<location*>(malloc(sizeof(location)));
  loc->x = static_cast<double*>(calloc(n > 0 ? n : 1, sizeof(double)));
  loc->y = static_cast<double*>(calloc(n > 0 ? n : 1, sizeof(double)));
  loc->z = static_cast<double*>(calloc(n > 0 ? n : 1, sizeof(double)));
  for (int i = 0; i < n && i < (int)lines.size(); ++i) {
    char* save = nullptr;
    std::vector<char> buf(lines[i].begin(), lines[i].end());
    buf.push_back('\0');
    const char* a = strtok_r(buf.data(), ",", &save);
    const char* b = a ? strtok_r(nullptr, ",", &save) : nullptr;
    const char* c = b ? strtok_r(nullptr, ",", &save) : nullptr;
    loc->x[i] = a ? atof(a) : 0.0;
    loc->y[i] = b ? atof(b) : 0.0;
    loc->z[i] = c ? atof(c) : 0.0;
  }
  return loc;
}

// flat_file.c:115-145 (readObsFile) and 147-175 (readTimeFile): one value
// per line.
static double* read_values(const char* path, int n, const char* who) {
  std::vector<std::string> lines;
  if (!read_lines(path, lines)) {
    printf("%s:cannot open file: %s\n", who, path);
    return nullptr;
  }
  auto* v = static_cast<double*>(calloc(n > 0 ? n : 1, sizeof(double)));
  for (int i = 0; i < n && i < (int)lines.size(); ++i) v[i] = atof(lines[i].c_str());
  return v;
}

double* readObsFile(char* obsfile, int n) { return read_values(obsfile, n, "readObsFile"); }

double* readTimeFile(char* timefile, int n) { return read_values(timefile, n, "readTimeFile"); }

// --------------------------------------------------------- R .C wrappers
// r-wrappers/compute/rwrappers.c:1172-1187.
void rexageostat_init(int* ncores, int* gpus, int* dts, int* lts) { exageostat_init(ncores, gpus, dts, lts); }

void rexageostat_finalize(void) { exageostat_finalize(); }

// rwrappers.c:56-176: z = L(theta) e at the caller's locations, e from
// dlarnv(3, {0, 0, 0, 1}); globalvec (veclen >= rows) receives z.
void gen_z_givenlocs_exact(double* x, int* xlen, double* y, int* ylen, int* kernel, double* theta, int* thetalen,
                           int* dmetric, int* n, int* ncores, int* gpus, int* ts, int* p_grid, int* q_grid,
                           int* veclen, double* globalvec) {
  (void)xlen;
  (void)ylen;
  (void)thetalen;
  for (int j = 0; j < *veclen; ++j) globalvec[j] = 0.0;
  if (!check_kernel(*kernel, __func__)) return;
  const int rows = *n * kernel_vars(*kernel);
  if (*veclen < rows) {
    fprintf(stderr, "%s: veclen %d < %d rows\n", __func__, *veclen, rows);
    return;
  }
  MLE_data data;
  init_data_values(&data);
  data.kernel_fun = (char*)kernel_name(*kernel);
  data.dm = (char*)(*dmetric == 1 ? "gc" : "ed");
  data.l1.x = x;
  data.l1.y = y;
  std::vector<double> e(rows);
  int iseed[4] = {0, 0, 0, 1};
  exageostat_dlarnv(3, iseed, rows, e.data());
  EXAGEOSTAT_dmle_Call(&data, *ncores, *gpus, *ts, *p_grid, *q_grid, rows, rows, 0);
  if (exageostat_gpu_last_status() == 0) {
    MLE_zvg(&data, e.data(), theta, rows, *ts, 0, *p_grid, *q_grid);
    if (exageostat_gpu_last_status() == 0) MLE_get_zobs(&data, globalvec, rows);
  }
  MLE_Finalize(&data);
}

// rwrappers.c:178-301: GenerateXYLoc(n, seed), e from dlarnv(3, {seed,
// seed, seed, 1}), z = L(theta) e; globalvec = [x (n), y (n), z (rows)].
void gen_z_exact(int* kernel, double* theta, int* thetalen, int* dmetric, int* n, int* seed, int* ncores, int* gpus,
                 int* ts, int* p_grid, int* q_grid, int* veclen, double* globalvec) {
  (void)thetalen;
  for (int j = 0; j < *veclen; ++j) globalvec[j] = 0.0;
  if (!check_kernel(*kernel, __func__)) return;
  const int rows = *n * kernel_vars(*kernel);
  if (*veclen < 2 * *n + rows) {
    fprintf(stderr, "%s: veclen %d < %d\n", __func__, *veclen, 2 * *n + rows);
    return;
  }
  int iseed[4] = {*seed, *seed, *seed, 1};
  std::vector<double> e(rows);
  exageostat_dlarnv(3, iseed, rows, e.data());
  location* loc = GenerateXYLoc(*n, *seed);
  std::copy(loc->x, loc->x + *n, globalvec);
  std::copy(loc->y, loc->y + *n, globalvec + *n);
  MLE_data data;
  init_data_values(&data);
  data.kernel_fun = (char*)kernel_name(*kernel);
  data.dm = (char*)(*dmetric == 1 ? "gc" : "ed");
  data.l1.x = globalvec;
  data.l1.y = globalvec + *n;
  EXAGEOSTAT_dmle_Call(&data, *ncores, *gpus, *ts, *p_grid, *q_grid, rows, rows, 0);
  if (exageostat_gpu_last_status() == 0) {
    MLE_zvg(&data, e.data(), theta, rows, *ts, 0, *p_grid, *q_grid);
    if (exageostat_gpu_last_status() == 0) MLE_get_zobs(&data, globalvec + 2 * *n, rows);
  }
  MLE_Finalize(&data);
  free(loc->x);
  free(loc->y);
  free(loc);
}

// mle_general (rwrappers.c:508-707): BOBYQA over MLE_alg from the lower
// bounds; globalthetaout = [theta hat (np), seconds per evaluation, total
// seconds, evaluations].  comp_mode 0: exact; 2: diag_approx with band
// diag_thick; 3: mixed precision with FP64 band diag_thick.
static void mle_general(int kernel, double* x, double* y, double* z, double* clb, double* cub, int comp_mode,
                        int diag_thick, int dmetric, int n, double opt_tol, int opt_max_iters, int ncores, int gpus,
                        int ts, int p_grid, int q_grid, double* globalthetaout, const char* fn) {
  const int np = kernel_params(kernel);
  if (np > 0)
    for (int j = 0; j < np; ++j) globalthetaout[j] = -1;
  if (!check_kernel(kernel, fn)) return;
  if (kernel == 5) {
    fprintf(stderr, "%s: the space-time kernel needs time coordinates, which .C gives no argument for\n", fn);
    return;
  }
  const int rows = n * kernel_vars(kernel);
  MLE_data data;
  init_data_values(&data);
  data.kernel_fun = (char*)kernel_name(kernel);
  data.dm = (char*)(dmetric == 1 ? "gc" : "ed");
  data.l1.x = x;
  data.l1.y = y;
  data.precision = 0;
  if (comp_mode == 2) {
    data.computation = (char*)"diag_approx";
    data.diag_thick = diag_thick;
  } else if (comp_mode == 3) {
    data.precision = 2;
    data.diag_thick = diag_thick > 0 ? diag_thick : 1;
  }
  EXAGEOSTAT_dmle_Call(&data, ncores, gpus, ts, p_grid, q_grid, rows, 0, 0);
  if (exageostat_gpu_last_status() != 0) {
    MLE_Finalize(&data);
    return;
  }
  EXAGEOSTAT_Lapack_to_Tile(z, rows, static_cast<CHAM_desc_t*>(data.descZ));
  std::vector<double> theta(clb, clb + np);
  double fmax = 0.0;
  int nevals = 0;
  const double t0 = now();
  const int rc = exageostat_bobyqa((unsigned)np, MLE_alg, &data, clb, cub, theta.data(), opt_tol, opt_max_iters,
                                   &fmax, &nevals);
  data.total_exec_time = now() - t0;
  MLE_Finalize(&data);
  if (rc < 0) return;
  for (int j = 0; j < np; ++j) globalthetaout[j] = theta[j];
  globalthetaout[np] = data.iter_count > 0 ? data.total_exec_time / (double)data.iter_count : 0.0;
  globalthetaout[np + 1] = data.total_exec_time;
  globalthetaout[np + 2] = (double)data.iter_count;
}

// rwrappers.c:710-753.
void mle_exact(double* x, int* xlen, double* y, int* ylen, double* z, int* zlen, double* clb, int* clblen, double* cub,
               int* cublen, int* kernel, int* dmetric, int* n, double* opt_tol, int* opt_max_iters, int* ncores,
               int* gpus, int* ts, int* p_grid, int* q_grid, double* globalthetaout) {
  (void)xlen, (void)ylen, (void)zlen, (void)clblen, (void)cublen;
  mle_general(*kernel, x, y, z, clb, cub, 0, 0, *dmetric, *n, *opt_tol, *opt_max_iters, *ncores, *gpus, *ts,
              *p_grid, *q_grid, globalthetaout, __func__);
}

// rwrappers.c:848-891.
void mle_dst(double* x, int* xlen, double* y, int* ylen, double* z, int* zlen, int* kernel, double* clb, int* clblen,
             double* cub, int* cublen, int* dst_thick, int* dmetric, int* n, double* opt_tol, int* opt_max_iters,
             int* ncores, int* gpus, int* ts, int* p_grid, int* q_grid, double* globalthetaout) {
  (void)xlen, (void)ylen, (void)zlen, (void)clblen, (void)cublen;
  mle_general(*kernel, x, y, z, clb, cub, 2, *dst_thick, *dmetric, *n, *opt_tol, *opt_max_iters, *ncores, *gpus,
              *ts, *p_grid, *q_grid, globalthetaout, __func__);
}

// rwrappers.c:893-936.
void mle_mp(double* x, int* xlen, double* y, int* ylen, double* z, int* zlen, int* kernel, double* clb, int* clblen,
            double* cub, int* cublen, int* mp_band, int* dmetric, int* n, double* opt_tol, int* opt_max_iters,
            int* ncores, int* gpus, int* ts, int* p_grid, int* q_grid, double* globalthetaout) {
  (void)xlen, (void)ylen, (void)zlen, (void)clblen, (void)cublen;
  mle_general(*kernel, x, y, z, clb, cub, 3, *mp_band, *dmetric, *n, *opt_tol, *opt_max_iters, *ncores, *gpus, *ts,
              *p_grid, *q_grid, globalthetaout, __func__);
}

// rwrappers.c:938-1072: kriging at (xmiss, ymiss) from (xobs, yobs, zobs);
// globalzpredict receives the predictions (-1000 where nothing was
// computed).  nZobs / nZmiss count locations; bivariate kernels return two
// interleaved values per held-out location.
void exact_predict(double* xobs, int* xobs_len, double* yobs, int* yobs_len, double* zobs, int* zobs_len,
                   double* xmiss, int* xmiss_len, double* ymiss, int* ymiss_len, int* nZobs, int* nZmiss, int* kernel,
                   double* est_theta, int* thetalen, int* computation, int* dmetric, int* ncores, int* gpus, int* ts,
                   int* p_grid, int* q_grid, double* globalzpredict) {
  (void)xobs_len, (void)yobs_len, (void)zobs_len, (void)xmiss_len, (void)ymiss_len, (void)thetalen;
  if (!check_kernel(*kernel, __func__)) return;
  const int p = kernel_vars(*kernel);
  for (int j = 0; j < p * *nZmiss; ++j) globalzpredict[j] = -1000;
  if (*computation != 0) {
    printf("In R mode: only exact prediction is allowed!\n");
    return;
  }
  if (*kernel == 5) {
    fprintf(stderr, "%s: the space-time kernel needs time coordinates, which .C gives no argument for\n", __func__);
    return;
  }
  MLE_data data;
  init_data_values(&data);
  data.kernel_fun = (char*)kernel_name(*kernel);
  data.dm = (char*)(*dmetric == 1 ? "gc" : "ed");
  data.l1.x = xobs;
  data.l1.y = yobs;
  data.lobs.x = xobs;
  data.lobs.y = yobs;
  data.lmiss.x = xmiss;
  data.lmiss.y = ymiss;
  // The likelihood matrix is not used: the context holds one tile.
  EXAGEOSTAT_dmle_Call(&data, *ncores, *gpus, *ts, *p_grid, *q_grid, *ts, p * *nZobs, p * *nZmiss);
  if (exageostat_gpu_last_status() == 0) {
    prediction_init(&data, *nZmiss, *nZobs, *ts, *p_grid, *q_grid, 0);
    std::vector<double> zmiss(p * *nZmiss, 0.0);
    if (exageostat_gpu_last_status() == 0) {
      EXAGEOSTAT_dmle_Predict_Tile(&data, est_theta, p * *nZmiss, p * *nZobs, zobs, nullptr, zmiss.data(),
                                   p * *nZobs);
      if (exageostat_gpu_last_status() == 0) std::copy(zmiss.begin(), zmiss.end(), globalzpredict);
    }
    prediction_finalize(&data);
  }
  MLE_Finalize(&data);
}

// rwrappers.c:1074-1170: MLOE and MMOM of kriging with est_theta when
// true_theta holds; global_mloe_mmom = [MLOE, MMOM] (-1000 on failure).
void exact_mloe_mmom(double* xobs, int* xobs_len, double* yobs, int* yobs_len, double* zobs, int* zobs_len,
                     double* xmiss, int* xmiss_len, double* ymiss, int* ymiss_len, int* nZobs, int* nZmiss,
                     int* kernel, double* est_theta, double* true_theta, int* thetalen, int* dmetric, int* ncores,
                     int* gpus, int* ts, int* p_grid, int* q_grid, double* global_mloe_mmom) {
  (void)xobs_len, (void)yobs_len, (void)zobs, (void)zobs_len, (void)xmiss_len, (void)ymiss_len, (void)thetalen;
  global_mloe_mmom[0] = global_mloe_mmom[1] = -1000;
  if (!check_kernel(*kernel, __func__)) return;
  MLE_data data;
  init_data_values(&data);
  data.kernel_fun = (char*)kernel_name(*kernel);
  data.dm = (char*)(*dmetric == 1 ? "gc" : "ed");
  data.l1.x = xobs;
  data.l1.y = yobs;
  data.lobs.x = xobs;
  data.lobs.y = yobs;
  data.lmiss.x = xmiss;
  data.lmiss.y = ymiss;
  EXAGEOSTAT_dmle_Call(&data, *ncores, *gpus, *ts, *p_grid, *q_grid, *ts, *nZobs, *nZmiss);
  if (exageostat_gpu_last_status() == 0) {
    mloe_mmom_init(&data, *nZmiss, *nZobs, *ts, *p_grid, *q_grid);
    EXAGEOSTAT_dmle_mloe_mmom_Tile(&data, true_theta, est_theta, *nZmiss, *nZobs, *nZobs + *nZmiss);
    if (exageostat_gpu_last_status() == 0 && data.status == 0) {
      global_mloe_mmom[0] = data.mloe;
      global_mloe_mmom[1] = data.mmom;
    }
    mloe_mmom_finalize(&data);
  }
  MLE_Finalize(&data);
}

}  // extern "C"
