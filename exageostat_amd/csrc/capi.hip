// The reference's C entry points for the exact-dense likelihood path,
// implemented over the HIP engine.  Declarations and citations are in
// include/exageostat_gpu.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/exageostat_gpu.h"
#include "engine.hpp"
#include "mixed.hpp"
#include "adjoint.hpp"
#include "mloe.hpp"
#include "profile.hpp"

using namespace exa;

namespace {

thread_local int g_status = 0;
thread_local std::string g_error;

void set_error(int status, const std::string& msg) {
  g_status = status;
  g_error = msg;
}

// Library-wide streams and workspace for the descriptor-level calls.
struct Runtime {
  hipStream_t s1 = nullptr, s2 = nullptr;
  CholWorkspace ws;
  double* d_partial = nullptr;
  long partial_len = 0;
  Runtime() {
    EXA_HIP(hipStreamCreateWithFlags(&s1, hipStreamNonBlocking));
    EXA_HIP(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
  }
  double* partial(long n) {
    if (n > partial_len) {
      if (d_partial) (void)hipFree(d_partial);
      EXA_HIP(hipMalloc(&d_partial, sizeof(double) * n));
      partial_len = n;
    }
    return d_partial;
  }
};
Runtime* g_rt = nullptr;

Runtime& rt() {
  if (!g_rt) g_rt = new Runtime();
  return *g_rt;
}

// Descriptor payload: a device tile matrix (owned or borrowed), or caller
// host memory for scalar accumulators.
struct DescImpl {
  TileMatrix* tm = nullptr;
  DeviceVector* vec = nullptr;  // column vectors (Z, Zcpy) are plain padded arrays
  bool owns = false;
  double* host = nullptr;
  DeviceLocations rows, cols;  // location cache for dcmg on this matrix
  ~DescImpl() {
    if (owns) delete tm;
  }
};

// The payload of a descriptor: std::invalid_argument for a null descriptor
// and for a tile-matrix descriptor whose matrix is not allocated (the FP64
// covariance of a lean context).
DescImpl* impl_of(const CHAM_desc_t* d) {
  if (!d || !d->mat) throw std::invalid_argument("null descriptor");
  auto* impl = static_cast<DescImpl*>(d->mat);
  if (impl->tm && !impl->tm->allocated())
    throw std::invalid_argument("the descriptor's tile matrix is not allocated (descC of a lean context)");
  return impl;
}

// The tile matrix of a descriptor payload (not a vector or scalar one).
TileMatrix& need_tm(DescImpl* impl) {
  if (!impl->tm) throw std::invalid_argument("the descriptor is not a tile matrix");
  return *impl->tm;
}

CHAM_desc_t* make_desc(TileMatrix* tm, bool owns) {
  auto* d = new CHAM_desc_t();
  auto* impl = new DescImpl();
  impl->tm = tm;
  impl->owns = owns;
  d->mat = impl;
  d->dtyp = ChamRealDouble;
  d->mb = d->nb = tm->ts;
  d->bsiz = tm->ts * tm->ts;
  d->m = tm->m;
  d->n = tm->n;
  d->mt = tm->mt;
  d->nt = tm->nt;
  d->lm = tm->mt * tm->ts;
  d->ln = tm->nt * tm->ts;
  d->uplo = tm->lower ? ChamLower : ChamUpperLower;
  return d;
}

CHAM_desc_t* make_vec_desc(DeviceVector* v, int ts) {
  auto* d = new CHAM_desc_t();
  auto* impl = new DescImpl();
  impl->vec = v;
  d->mat = impl;
  d->dtyp = ChamRealDouble;
  d->mb = ts;
  d->nb = 1;
  d->bsiz = ts;
  d->m = (int)v->n;
  d->n = 1;
  d->mt = (int)(v->padded / ts);
  d->nt = 1;
  d->lm = (int)v->padded;
  d->ln = 1;
  d->uplo = ChamUpperLower;
  return d;
}

CHAM_desc_t* make_user_desc(double* host, int m, int n) {
  auto* d = new CHAM_desc_t();
  auto* impl = new DescImpl();
  impl->host = host;
  d->mat = impl;
  d->dtyp = ChamRealDouble;
  d->mb = d->nb = 1;
  d->bsiz = 1;
  d->m = d->lm = m;
  d->n = d->ln = n;
  d->mt = m;
  d->nt = n;
  d->uplo = ChamUpperLower;
  return d;
}

void free_desc(void*& p) {
  if (!p) return;
  auto* d = static_cast<CHAM_desc_t*>(p);
  delete static_cast<DescImpl*>(d->mat);
  delete d;
  p = nullptr;
}

// Sequence = stream + host reductions to finish at wait time.
struct SeqImpl {
  hipStream_t s = nullptr;
  bool owns = false;
  struct Pending {
    std::vector<double> h;
    double* target;
  };
  std::vector<Pending> pending;
};

hipStream_t seq_stream(RUNTIME_sequence_t* seq) {
  if (seq && seq->impl) return static_cast<SeqImpl*>(seq->impl)->s;
  return rt().s1;
}

// Scratch of EXAGEOSTAT_dtrsv_Tile_noalloc, the entry point that allocates
// nothing: the padded right-hand side (kNoAllocRows doubles), the solve's
// assembly vector (kNoAllocRows), the partial sums of the diagonal-tile
// products ((kNoAllocTs / 64) * kNoAllocTs) and the completion ticket.
// 1.13 MB per GPU, held until the process ends.  Host code never passes
// `scratch` itself (its host-side name is not a device address): it uses
// the device address that hipGetSymbolAddress returns for the current
// device.
constexpr long kNoAllocRows = 65536;  // mt * ts at most
constexpr long kNoAllocTs = 1024;
constexpr long kNoAllocPart = (kNoAllocTs / kNB) * kNoAllocTs;
constexpr long kNoAllocDoubles = 2 * kNoAllocRows + kNoAllocPart + 32;
static __device__ __align__(256) double scratch[kNoAllocDoubles];

// Device address of `scratch` on the current device, looked up once per
// device.
double* scratch_ptr() {
  static std::vector<double*> ptrs;
  int dev = 0;
  EXA_HIP(hipGetDevice(&dev));
  if ((int)ptrs.size() <= dev) ptrs.resize(dev + 1, nullptr);
  if (!ptrs[dev]) {
    void* p = nullptr;
    const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (rc != hipSuccess || !p)
      throw std::runtime_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(rc));
    ptrs[dev] = static_cast<double*>(p);
  }
  return ptrs[dev];
}

Context* ctx_of(MLE_data* data) {
  if (!data || !data->gpu) throw std::runtime_error("MLE_data has no GPU context (call EXAGEOSTAT_dmle_Call)");
  return static_cast<Context*>(data->gpu);
}

int metric_of(const char* dm) { return (dm && !strcmp(dm, "gc")) ? 1 : 0; }

bool is_bivariate(const char* k) {
  return k && (!strncmp(k, "bivariate_matern_parsimonious", 29));
}

void copy_vec_to_device(DeviceVector& v, const double* h, long n, hipStream_t s) {
  EXA_HIP(hipMemsetAsync(v.d, 0, sizeof(double) * v.padded, s));
  EXA_HIP(hipMemcpyAsync(v.d, h, sizeof(double) * n, hipMemcpyHostToDevice, s));
  EXA_HIP(hipStreamSynchronize(s));
}

// The reference's per-iteration lines (MLE_dexact.c:439-472), also to
// data->pFileLog when data->log == 1.
void print_iteration(MLE_data* data, const double* theta, int np, double loglik) {
  FILE* logf = (data->log == 1) ? data->pFileLog : nullptr;
  if (data->silent && !logf) return;
  for (FILE* f : {data->silent ? (FILE*)nullptr : stderr, logf}) {
    if (!f) continue;
    fprintf(f, " %3d- Model Parameters (", data->iter_count + 1);
    for (int i = 0; i < np; ++i) fprintf(f, "%.8f%s", theta[i], i + 1 < np ? "," : "");
    fprintf(f, ")----> LogLi: %.18f\n", loglik);
    if (f == logf) continue;
    fprintf(f, " ---- Facto Time: %6.2f\n", data->time_facto);
    fprintf(f, " ---- Matrix Generation Time: %6.2f\n", data->time_gen);
    fprintf(f, " ---- Total Time: %6.2f\n", data->time_total);
  }
}

// recover() of misc/compute/MLE_misc.c:2035-2070: the line of `path` whose
// first field is iter_count holds theta (np values) and the likelihood.
// Returns true and fills theta_out / loglik when the line exists.
bool recover_iteration(const char* path, int iter_count, int np, std::vector<double>& theta_out, double& loglik) {
  FILE* fp = fopen(path, "r");
  if (!fp) throw std::runtime_error(std::string("cannot open the recovery file ") + path);
  char* line = nullptr;
  size_t cap = 0;
  bool found = false;
  while (!found && getline(&line, &cap, fp) != -1) {
    char* save = nullptr;
    char* tok = strtok_r(line, " \t\n", &save);
    if (!tok || atoi(tok) != iter_count) continue;
    theta_out.assign(np, 0.0);
    int got = 0;
    for (; got < np && (tok = strtok_r(nullptr, " \t\n", &save)); ++got) theta_out[got] = atof(tok);
    tok = strtok_r(nullptr, " \t\n", &save);
    if (got == np && tok) {
      loglik = atof(tok);
      found = true;
    }
  }
  free(line);
  fclose(fp);
  return found;
}

// checkpointing() of misc/compute/MLE_misc.c:2014-2033: one line
// "iter theta... loglik" appended per evaluation (the format recover reads).
void checkpoint_iteration(const char* path, int iter_count, const double* theta, int np, double loglik) {
  FILE* fp = fopen(path, "a");
  if (!fp) throw std::runtime_error(std::string("cannot open the checkpoint file ") + path);
  fprintf(fp, "%d ", iter_count);
  for (int i = 0; i < np; ++i) fprintf(fp, "%.17g ", theta[i]);
  fprintf(fp, "%.17g \n", loglik);
  fclose(fp);
}

// Morton order of locations_obs_zsort_inplace (misc/compute/MLE_misc.c:
// 405-535): coordinates rescaled to [0, 2^32-1], keys interleave y above x
// bit by bit; ties keep (even popcount) or reverse (odd) their input order,
// which is what the reference's in-place partition sort produces.
void obs_zsort(long n, double* x, double* y, double* z) {
  if (n <= 1) return;
  double mn[2] = {x[0], y[0]}, mx[2] = {x[0], y[0]};
  for (long i = 1; i < n; ++i) {
    mn[0] = std::min(mn[0], x[i]);
    mx[0] = std::max(mx[0], x[i]);
    mn[1] = std::min(mn[1], y[i]);
    mx[1] = std::max(mx[1], y[i]);
  }
  struct Item {
    uint64_t key;
    long tie;
    long idx;
  };
  std::vector<Item> it(n);
  for (long i = 0; i < n; ++i) {
    const uint32_t ux = (uint32_t)((x[i] - mn[0]) / (mx[0] - mn[0]) * 4294967295.0);
    const uint32_t uy = (uint32_t)((y[i] - mn[1]) / (mx[1] - mn[1]) * 4294967295.0);
    uint64_t key = 0;
    for (int b = 31; b >= 0; --b) {
      key = (key << 1) | ((uy >> b) & 1u);
      key = (key << 1) | ((ux >> b) & 1u);
    }
    const int pop = __builtin_popcountll(key);
    it[i] = {key, (pop % 2 == 0) ? i : -i, i};
  }
  std::sort(it.begin(), it.end(), [](const Item& a, const Item& b) {
    return a.key != b.key ? a.key < b.key : a.tie < b.tie;
  });
  std::vector<double> nx(n), ny(n), nz(n);
  for (long i = 0; i < n; ++i) {
    nx[i] = x[it[i].idx];
    ny[i] = y[it[i].idx];
    nz[i] = z[it[i].idx];
  }
  std::copy(nx.begin(), nx.end(), x);
  std::copy(ny.begin(), ny.end(), y);
  std::copy(nz.begin(), nz.end(), z);
}

uint32_t part1by1(uint32_t v) {
  v &= 0x0000ffffu;
  v = (v | (v << 8)) & 0x00ff00ffu;
  v = (v | (v << 4)) & 0x0f0f0f0fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}

uint32_t compact1by1(uint32_t v) {
  v &= 0x55555555u;
  v = (v | (v >> 1)) & 0x33333333u;
  v = (v | (v >> 2)) & 0x0f0f0f0fu;
  v = (v | (v >> 4)) & 0x00ff00ffu;
  v = (v | (v >> 8)) & 0x0000ffffu;
  return v;
}

template <class F>
auto guarded(F&& f, decltype(f()) on_error) -> decltype(f()) {
  try {
    g_status = 0;
    g_error.clear();
    return f();
  } catch (const std::invalid_argument& e) {
    set_error(-1, e.what());
  } catch (const std::exception& e) {
    set_error(-2, e.what());
  }
  return on_error;
}

void release_prediction_descs(MLE_data* data) {
  free_desc(data->descC22);
  free_desc(data->descC12);
  free_desc(data->descZobs);
  free_desc(data->descZmiss);
  free_desc(data->descZactual);
  free_desc(data->descmse);
  free_desc(data->descmse1);
  free_desc(data->descmse2);
}

// Kriging (MLE_dexact.c:688-874): C22 = Sigma(lobs, lobs) and C12 =
// Sigma(lmiss, lobs) generated on the device, C22 factored, alpha =
// C22^-1 zobs by two triangular solves, zmiss = C12 alpha into ctx->Zmiss.
// zobs: host values, or null to use ctx->Zobs as the caller filled it.
// The mixed-precision options of MLE_data (single: every off-diagonal tile
// FP32, no refinement, as in the FP32 likelihood path).
MixedOptions mixed_options_of(const MLE_data* data, bool single) {
  MixedOptions opt;
  if (single) {
    opt.ir_steps = 0;
    return opt;
  }
  opt.fp64_corr = data->mp_fp64_corr;
  opt.fp32_corr = data->mp_fp32_corr;
  opt.bf16 = data->mp_bf16 != 0;
  opt.ir_steps = data->mp_ir_steps;
  opt.ir_tol = data->mp_ir_tol;
  opt.diag_thick = data->diag_thick;
  opt.lolo = data->mp_lolo != 0;
  opt.grad_f32 = data->mp_grad_f32;
  return opt;
}

// Kriging Zmiss = C12 C22^-1 Zobs (MLE_dexact.c:688-874).  With mopt, the
// single- or mixed-precision variant (MLE_sexact.c, MLE_sdexact.c:528-680:
// the reference factors C22 in its lower precision and solves with that
// factor): C22 is factored by the mixed-precision engine under mopt (per-
// theta tile precisions, FP64 refinement of the solve when mopt asks for
// it) on a context of the observed locations; C12 and the product stay
// FP64.  The refinement's non-convergence is reported as in the
// likelihood (status -3, the predictions still returned).
void predict_core(MLE_data* data, Context* ctx, KernelId kid, const double* theta, int nZmiss, int nZobs,
                  const location& lmiss, const location& lobs, const double* zobs,
                  const MixedOptions* mopt = nullptr) {
  const int vars = kernel_vars(kid);
  if (!ctx->C12.allocated() || ctx->C12.m != nZmiss || ctx->C12.n != nZobs ||
      (!mopt && (!ctx->C22.allocated() || ctx->C22.m != nZobs))) {
    EXAGEOSTAT_dmle_Predict_Allocate(data, nZmiss / vars, nZobs / vars, ctx->ts, 1, 1, 1);
    if (g_status < 0) throw std::runtime_error(g_error);
  }
  const long nobs_loc = nZobs / vars;
  const long nmiss_loc = nZmiss / vars;
  hipStream_t s = ctx->s1;
  ctx->lobs.upload(lobs.x, lobs.y, nobs_loc, s, lobs.z);
  ctx->lmiss.upload(lmiss.x, lmiss.y, nmiss_loc, s, lmiss.z);
  const double t0 = wall_seconds();
  const int metric = metric_of(data->dm);
  generate_covariance(ctx->C12, kid, theta, ctx->lmiss, ctx->lobs, metric, s);
  if (zobs) copy_vec_to_device(ctx->Zobs, zobs, nZobs, s);
  if (mopt) {
    if (!ctx->pred || ctx->pred->N != nZobs || ctx->pred->ts != ctx->ts || ctx->pred->lean != ctx->lean) {
      delete ctx->pred;
      ctx->pred = nullptr;
      ctx->pred = new Context(nZobs, ctx->ts, 0, 0, ctx->lean);
    }
    Context& po = *ctx->pred;
    EXA_HIP(hipStreamSynchronize(s));
    po.l1.upload(lobs.x, lobs.y, nobs_loc, po.s1, lobs.z);
    EXA_HIP(hipMemcpyAsync(po.Zcpy.d, ctx->Zobs.d, sizeof(double) * nZobs, hipMemcpyDeviceToDevice, po.s1));
    double logdet = 0.0, dotp = 0.0;
    int info = 0;
    MixedReport rep;
    mixed_loglik(po, kid, theta, metric, *mopt, &logdet, &dotp, &info, &rep);
    if (info != 0) throw std::runtime_error("prediction: C22 is not positive definite");
    data->mp_ir_done = rep.ir_done;
    data->mp_ir_change = rep.ir_change;
    if (!rep.ir_converged) {
      set_error(-3, "mixed precision: the iterative refinement of the prediction's solve did not reach mp_ir_tol in "
                    "mp_ir_steps steps");
      data->status = -3;
    }
    // Zobs then holds C22^-1 Zobs, as after the FP64 path's two solves
    EXA_HIP(hipMemcpyAsync(ctx->Zobs.d, po.work.d, sizeof(double) * nZobs, hipMemcpyDeviceToDevice, po.s1));
    EXA_HIP(hipStreamSynchronize(po.s1));
  } else {
    generate_covariance(ctx->C22, kid, theta, ctx->lobs, ctx->lobs, metric, s);
    const int info = tile_potrf(ctx->C22, ctx->ws, s, ctx->s2, true);
    if (info != 0) throw std::runtime_error("prediction: C22 is not positive definite");
    tile_trsv(ctx->C22, ctx->ws, ctx->Zobs.d, 0, s);
    tile_trsv(ctx->C22, ctx->ws, ctx->Zobs.d, 1, s);
  }
  launch_gemv_full(ctx->C12.view(), ctx->Zobs.d, ctx->Zmiss.d, ctx->gemv_part.d, s);
  EXA_HIP(hipStreamSynchronize(s));
  const double t1 = wall_seconds();
  ctx->last.total = t1 - t0;
  data->time_total = t1 - t0;
  results.total_pred_time = t1 - t0;
}

// MSE (codelet_zmse.c:22-38) and, for the bivariate kernels, per variable
// (codelet_zmse_bivariate.c:22-51: by row parity; the block layout of
// "bivariate_matern_parsimonious2" splits the rows into halves instead).
double prediction_mse(MLE_data* data, KernelId kid, int vars, const double* Zactual, const double* Zmiss,
                      int nZmiss) {
  double se = 0.0, se1 = 0.0, se2 = 0.0;
  for (int i = 0; i < nZmiss; ++i) {
    const double e = (Zactual[i] - Zmiss[i]) * (Zactual[i] - Zmiss[i]);
    se += e;
    const bool second = kid == kBivariateBlock ? (i >= nZmiss / 2) : (i % 2 == 1);
    if (second) se2 += e; else se1 += e;
  }
  data->mserror = se / nZmiss;
  data->mserror1 = data->mserror2 = 0.0;
  if (vars == 2) {
    data->mserror1 = se1 / (nZmiss / 2);
    data->mserror2 = se2 / (nZmiss / 2);
  }
  results.mse_pred = data->mserror;
  results.mse_pred1 = data->mserror1;
  results.mse_pred2 = data->mserror2;
  return data->mserror;
}

// Likelihood paths of MLE_alg (src/compute/MLE.c:138-193).
enum class Path {
  kExact,   // precision 0, "exact": EXAGEOSTAT_dmle_Tile (MLE_dexact.c:231-489)
  kMixed,   // precision 2: EXAGEOSTAT_sdmle_Tile (MLE_sdexact.c:172-357)
  kSingle,  // precision 1: EXAGEOSTAT_smle_Tile (MLE_sexact.c:155-277)
  kBand,    // "diag_approx": EXAGEOSTAT_dmle_diag_Tile (MLE_approx.c:201-365)
};

// One likelihood evaluation, shared by the entry points of every path: the
// bookkeeping of the reference's evaluation functions (z restore, recovery
// replay, iteration line, averages, results) around the engine call.
double evaluate(MLE_data* data, unsigned n, const double* theta, double* grad, Path path) {
  const double out = guarded([&]() {
    Context* ctx = ctx_of(data);
    const KernelId kid = kernel_from_name(data->kernel_fun);
    if (kid == kUnknownKernel)
      throw std::invalid_argument(std::string("unsupported kernel: ") + (data->kernel_fun ? data->kernel_fun : "(null)"));
    const bool profile = kernel_is_profile(data->kernel_fun);
    const int np = n ? (int)n : kernel_num_params(kid);
    const int metric = metric_of(data->dm);
    const double* lz = (kid == kSpaceTime || data->l1.z) ? data->l1.z : nullptr;
    if (grad) {
      // The single-precision factor (no refinement) and the band-limited one
      // carry no Sigma^-1 accurate enough for the trace term; a gradient-
      // based optimiser must not iterate on a made-up value, so those calls
      // fail and say why.
      if (path != Path::kExact && path != Path::kMixed)
        throw std::invalid_argument(
            "only precision 0 with computation \"exact\" and precision 2 compute a gradient: pass grad = NULL "
            "for precision 1 and for diag_approx");
      if (!(kid == kUnivariate || kid == kNugget || kid == kBivariate) || profile || lz)
        throw std::invalid_argument(
            "the adjoint gradient covers the 2-D univariate Matern, nugget and interleaved parsimonious bivariate "
            "kernels");
    }
    const long nloc = ctx->N / kernel_vars(kid);
    ctx->l1.upload(data->l1.x, data->l1.y, nloc, ctx->s1, lz);
    if (data->iter_count == 0) {
      // Keep a pristine copy of Z for later iterations (MLE_dexact.c:303-317).
      EXA_HIP(hipMemcpyAsync(ctx->Zcpy.d, ctx->Z.d, sizeof(double) * ctx->Z.padded, hipMemcpyDeviceToDevice,
                             ctx->s1));
      EXA_HIP(hipStreamSynchronize(ctx->s1));
    }
    double loglik = 0.0;
    bool replayed = false;
    std::vector<double> rtheta;
    if (!grad && data->recovery_file && data->recovery_file[0])
      replayed = recover_iteration(data->recovery_file, data->iter_count, np, rtheta, loglik);
    int info = 0;
    if (!replayed) {
      double logdet = 0.0, dotp = 0.0;
      MixedReport rep;
      if (path == Path::kExact || path == Path::kBand) {
        const int band = path == Path::kBand ? std::max(1, data->diag_thick) : 0;
        if (grad) loglik = adjoint_loglik_grad(*ctx, kid, theta, metric, grad, &logdet, &dotp, &info);
        else loglik = ctx->loglik(kid, theta, metric, &logdet, &dotp, &info, band);
      } else {
        const MixedOptions opt = mixed_options_of(data, path == Path::kSingle);
        if (grad) loglik = mixed_adjoint_loglik_grad(*ctx, kid, theta, metric, opt, grad, &logdet, &dotp, &info, &rep);
        else loglik = mixed_loglik(*ctx, kid, theta, metric, opt, &logdet, &dotp, &info, &rep);
        data->mp_ir_done = rep.ir_done;
        data->mp_ir_change = rep.ir_change;
        data->mp_grad_check = rep.sweep_residual;
        if (path == Path::kSingle) data->sdotp = (float)dotp;
      }
      const long N = ctx->N;
      if (profile && info == 0) {
        // Profile likelihood of the "_profile" kernels (MLE_dexact.c:403-421),
        // N/2 in integer arithmetic as there; variance_v = |(L^-1 z)_v|^2 / (N/2).
        const long half = N / 2;
        // The band-limited path writes the interleaved kernel's profile with
        // (N/2.0) log(N/2.0) in place of (N/2) log N (MLE_approx.c:386-396).
        const bool band_form = path == Path::kBand && kid == kBivariate;
        if (band_form)
          loglik = -(N / 2.0) + (N / 2.0) * std::log(N / 2.0) - (N / 2.0) * std::log(dotp) - 0.5 * logdet -
                   (N / 2.0) * std::log(2.0 * M_PI);
        else
          loglik = -(double)half + (double)half * std::log((double)N) - (double)half * std::log(dotp) -
                   0.5 * logdet - (N / 2.0) * std::log(2.0 * M_PI);
        data->dotp1 = data->dotp2 = std::nan("");
        if (path == Path::kExact || path == Path::kBand) {
          std::vector<double> w(N);
          EXA_HIP(hipMemcpyAsync(w.data(), ctx->work.d, sizeof(double) * N, hipMemcpyDeviceToHost, ctx->s1));
          EXA_HIP(hipStreamSynchronize(ctx->s1));
          double d1 = 0.0, d2 = 0.0;
          for (long i = 0; i < N; ++i) {
            const bool second = (kid == kBivariateBlock) ? (i >= half) : (i & 1);
            (second ? d2 : d1) += w[i] * w[i];
          }
          data->dotp1 = d1;
          data->dotp2 = d2;
        }
        data->variance1 = data->dotp1 / (double)half;
        data->variance2 = data->dotp2 / (double)half;
      }
      data->det = 0.5 * logdet;
      data->dotp = dotp;
      data->logdet = logdet;
      data->time_gen = ctx->last.gen;
      data->time_facto = ctx->last.facto;
      data->time_logdet = ctx->last.logdet;
      data->time_solve = ctx->last.solve;
      data->time_total = ctx->last.total;
      data->status = info;
      if (info != 0) {
        set_error(info, "covariance matrix is not positive definite");
        loglik = -HUGE_VAL;
      } else if (!rep.ir_converged) {
        // l is returned, but the caller learns that the quadratic form did
        // not reach the requested accuracy.
        set_error(-3, "mixed precision: the iterative refinement did not reach mp_ir_tol in mp_ir_steps steps");
        data->status = -3;
      }
      data->avg_exec_time_per_iter += ctx->last.facto + ctx->last.logdet + ctx->last.solve;
      const double flops = (double)N * (((1.0 / 6.0) * N + 0.5) * N + 1.0 / 3.0) +
                           (double)N * (((1.0 / 6.0) * N) * N - 1.0 / 6.0) + (double)N * (N + 1.0);
      if (ctx->last.facto + ctx->last.solve > 0)
        data->avg_flops_per_iter += flops / 1e9 / (ctx->last.facto + ctx->last.solve);
      data->avg_generation_per_iter += ctx->last.gen;
      data->avg_cholesky_per_iter += ctx->last.facto;
      data->avg_solve_per_iter += ctx->last.solve;
    } else {
      data->status = 0;
    }
    print_iteration(data, theta, np, loglik);
    if (!replayed && data->checkpoint_file && data->checkpoint_file[0])
      checkpoint_iteration(data->checkpoint_file, data->iter_count, theta, np, loglik);
    data->variance = theta[0];
    data->iter_count++;
    data->final_loglik = loglik;
    results.final_loglik = loglik;
    if (results.estimated_theta) {
      int i = 0;
      if (profile) {
        results.estimated_theta[0] = data->variance1;
        results.estimated_theta[1] = data->variance2;
        i = 2;
      }
      for (; i < np; ++i) results.estimated_theta[i] = theta[i];
    }
    if (info != 0) g_status = info;
    else if (data->status == -3) g_status = -3;
    return loglik;
  }, -HUGE_VAL);
  if (g_status < 0 && g_status != -3 && data) data->status = g_status;
  return out;
}

}  // namespace

extern "C" {

output results;

int exageostat_gpu_last_status(void) { return g_status; }
const char* exageostat_gpu_last_error(void) { return g_error.c_str(); }

int exageostat_gpu_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

void exageostat_init(int* ncores, int* gpus, int* dts, int* lts) {
  (void)ncores;
  (void)gpus;
  (void)dts;
  (void)lts;
  guarded([&]() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) throw std::runtime_error("no HIP device");
    EXA_HIP(hipSetDevice(0));
    rt();
    return 0;
  }, 0);
}

void exageostat_finalize(void) {
  guarded([&]() {
    if (g_rt) EXA_HIP(hipDeviceSynchronize());
    return 0;
  }, 0);
}

// ------------------------------------------------------------ descriptors --
int EXAGEOSTAT_Desc_Create(CHAM_desc_t** desc, CHAM_enum uplo, int mb, int m, int n) {
  return guarded([&]() {
    auto* tm = new TileMatrix();
    try {
      tm->allocate(m, n, mb, uplo == ChamLower);
    } catch (...) {
      delete tm;
      throw;
    }
    *desc = make_desc(tm, true);
    return 0;
  }, -1);
}

int EXAGEOSTAT_Desc_Create_User(CHAM_desc_t** desc, double* host_mat, int m, int n) {
  return guarded([&]() {
    *desc = make_user_desc(host_mat, m, n);
    return 0;
  }, -1);
}

int EXAGEOSTAT_Desc_Destroy(CHAM_desc_t** desc) {
  if (!desc || !*desc) return 0;
  void* p = *desc;
  free_desc(p);
  *desc = nullptr;
  return 0;
}

int EXAGEOSTAT_Sequence_Create(RUNTIME_sequence_t** sequence) {
  return guarded([&]() {
    auto* s = new RUNTIME_sequence_t();
    auto* impl = new SeqImpl();
    EXA_HIP(hipStreamCreateWithFlags(&impl->s, hipStreamNonBlocking));
    impl->owns = true;
    s->impl = impl;
    s->status = 0;
    *sequence = s;
    return 0;
  }, -1);
}

int EXAGEOSTAT_Sequence_Wait(RUNTIME_sequence_t* sequence) {
  return guarded([&]() {
    if (!sequence || !sequence->impl) return 0;
    auto* impl = static_cast<SeqImpl*>(sequence->impl);
    EXA_HIP(hipStreamSynchronize(impl->s));
    for (auto& p : impl->pending) {
      double acc = 0.0;
      for (double v : p.h) acc += v;
      *p.target += acc;
    }
    impl->pending.clear();
    return sequence->status;
  }, -1);
}

int EXAGEOSTAT_Sequence_Destroy(RUNTIME_sequence_t* sequence) {
  if (!sequence) return 0;
  auto* impl = static_cast<SeqImpl*>(sequence->impl);
  if (impl) {
    if (impl->owns && impl->s) (void)hipStreamDestroy(impl->s);
    delete impl;
  }
  delete sequence;
  return 0;
}

int EXAGEOSTAT_Lapack_to_Tile(const double* A, int lda, CHAM_desc_t* desc) {
  return guarded([&]() {
    auto* impl = impl_of(desc);
    if (impl->host) {
      for (int j = 0; j < desc->n; ++j)
        for (int i = 0; i < desc->m; ++i) impl->host[i + (long)j * desc->m] = A[i + (long)j * lda];
      return 0;
    }
    if (impl->vec) {
      copy_vec_to_device(*impl->vec, A, impl->vec->n, rt().s1);
      return 0;
    }
    impl->tm->from_host(A, lda, rt().s1);
    return 0;
  }, -1);
}

int EXAGEOSTAT_Tile_to_Lapack(CHAM_desc_t* desc, double* A, int lda) {
  return guarded([&]() {
    auto* impl = impl_of(desc);
    if (impl->host) {
      for (int j = 0; j < desc->n; ++j)
        for (int i = 0; i < desc->m; ++i) A[i + (long)j * lda] = impl->host[i + (long)j * desc->m];
      return 0;
    }
    if (impl->vec) {
      EXA_HIP(hipMemcpyAsync(A, impl->vec->d, sizeof(double) * impl->vec->n, hipMemcpyDeviceToHost, rt().s1));
      EXA_HIP(hipStreamSynchronize(rt().s1));
      return 0;
    }
    impl->tm->to_host(A, lda, rt().s1);
    return 0;
  }, -1);
}

int EXAGEOSTAT_dpotrf_Tile(CHAM_enum uplo, CHAM_desc_t* A) {
  return guarded([&]() {
    if (uplo != ChamLower) throw std::invalid_argument("only ChamLower is supported");
    auto* impl = impl_of(A);
    TileMatrix& tm = need_tm(impl);
    if (!tm.lower) throw std::invalid_argument("descriptor must be lower-packed");
    return tile_potrf(tm, rt().ws, rt().s1, rt().s2, true);
  }, -1);
}

int EXAGEOSTAT_dpotri_Tile(CHAM_enum uplo, CHAM_desc_t* A) {
  return guarded([&]() {
    if (uplo != ChamLower) throw std::invalid_argument("only ChamLower is supported");
    auto* impl = impl_of(A);
    TileMatrix& tm = need_tm(impl);
    if (!tm.lower) throw std::invalid_argument("descriptor must be lower-packed");
    const size_t bytes = sizeof(double) * (size_t)tm.ts * tm.ts * tm.mt;
    double *P = nullptr, *Q = nullptr;
    EXA_HIP(hipMalloc(&P, bytes));
    if (hipMalloc(&Q, bytes) != hipSuccess) {
      (void)hipFree(P);
      throw std::runtime_error("potri: out of device memory");
    }
    try {
      tile_potri_inplace(tm, rt().ws, P, Q, rt().s1);
      EXA_HIP(hipStreamSynchronize(rt().s1));
    } catch (...) {
      (void)hipFree(P);
      (void)hipFree(Q);
      throw;
    }
    (void)hipFree(P);
    (void)hipFree(Q);
    return 0;
  }, -1);
}

int EXAGEOSTAT_dtrsv_Tile(CHAM_enum trans, CHAM_desc_t* L, double* b) {
  return guarded([&]() {
    auto* impl = impl_of(L);
    TileMatrix& tm = need_tm(impl);
    DeviceVector v;
    v.allocate(tm.m, (long)tm.mt * tm.ts);
    copy_vec_to_device(v, b, tm.m, rt().s1);
    tile_trsv(tm, rt().ws, v.d, trans == ChamTrans ? 1 : 0, rt().s1);
    EXA_HIP(hipMemcpyAsync(b, v.d, sizeof(double) * tm.m, hipMemcpyDeviceToHost, rt().s1));
    EXA_HIP(hipStreamSynchronize(rt().s1));
    return 0;
  }, -1);
}

int EXAGEOSTAT_dtrsv_Tile_noalloc(CHAM_enum trans, CHAM_desc_t* L, double* b) {
  return guarded([&]() {
    auto* impl = impl_of(L);
    if (!impl || !impl->tm) throw std::invalid_argument("L must be a tile-matrix descriptor");
    TileMatrix& tm = need_tm(impl);
    const long padded = (long)tm.mt * tm.ts;
    if (padded > kNoAllocRows || tm.ts > kNoAllocTs)
      throw std::invalid_argument("EXAGEOSTAT_dtrsv_Tile_noalloc: mt * ts = " + std::to_string(padded) +
                                  " and ts = " + std::to_string(tm.ts) + " must not exceed " +
                                  std::to_string(kNoAllocRows) + " and " + std::to_string(kNoAllocTs));
    if (rt().ws.factored != tm.tile_raw(0, 0) || rt().ws.mt_ < tm.mt || rt().ws.ts_ != tm.ts)
      throw std::invalid_argument("EXAGEOSTAT_dtrsv_Tile_noalloc: L is not the last matrix factored by "
                                  "EXAGEOSTAT_dpotrf_Tile");
    double* base = scratch_ptr();
    double* v = base;
    double* y = base + kNoAllocRows;
    double* tpart = y + kNoAllocRows;
    unsigned* ticket = reinterpret_cast<unsigned*>(tpart + kNoAllocPart);
    hipStream_t s = rt().s1;
    // the padding rows of v must be zero; the ticket starts at zero
    EXA_HIP(hipMemsetAsync(v, 0, sizeof(double) * padded, s));
    EXA_HIP(hipMemsetAsync(ticket, 0, sizeof(unsigned), s));
    EXA_HIP(hipMemcpyAsync(v, b, sizeof(double) * tm.m, hipMemcpyHostToDevice, s));
    tile_trsv_scratch(tm, rt().ws, v, trans == ChamTrans ? 1 : 0, s, y, tpart, ticket);
    EXA_HIP(hipMemcpyAsync(b, v, sizeof(double) * tm.m, hipMemcpyDeviceToHost, s));
    EXA_HIP(hipStreamSynchronize(s));
    return 0;
  }, -1);
}

// ------------------------------------------------------------- tile tasks --
int EXAGEOSTAT_MLE_dcmg_Tile_Async(CHAM_enum uplo, CHAM_desc_t* descA, location* l1, location* l2,
                                   location* lm, double* theta, char* dm, char* kernel_fun,
                                   RUNTIME_sequence_t* sequence, RUNTIME_request_t* request) {
  (void)lm;
  (void)request;
  (void)uplo;
  return guarded([&]() {
    if (sequence && sequence->status != 0) return -2;
    const KernelId kid = kernel_from_name(kernel_fun);
    if (kid == kUnknownKernel) throw std::invalid_argument(std::string("unsupported kernel: ") + (kernel_fun ? kernel_fun : "(null)"));
    auto* impl = impl_of(descA);
    TileMatrix& tm = need_tm(impl);
    const long nr = tm.m / kernel_vars(kid);
    const long nc = tm.n / kernel_vars(kid);
    hipStream_t s = seq_stream(sequence);
    impl->rows.upload(l1->x, l1->y, nr, s, l1->z);
    impl->cols.upload(l2->x, l2->y, nc, s, l2->z);
    generate_covariance(tm, kid, theta, impl->rows, impl->cols, metric_of(dm), s);
    return 0;
  }, -1);
}

int MORSE_MLE_dcmg_Tile_Async(CHAM_enum uplo, CHAM_desc_t* descA, location* l1, location* l2, location* lm,
                              double* theta, char* dm, char* kernel_fun, RUNTIME_sequence_t* sequence,
                              RUNTIME_request_t* request) {
  return EXAGEOSTAT_MLE_dcmg_Tile_Async(uplo, descA, l1, l2, lm, theta, dm, kernel_fun, sequence, request);
}

int EXAGEOSTAT_MLE_dmdet_Tile_Async(CHAM_desc_t* descA, RUNTIME_sequence_t* sequence,
                                    RUNTIME_request_t* request, CHAM_desc_t* descdet) {
  (void)request;
  return guarded([&]() {
    if (sequence && sequence->status != 0) return -2;
    auto* impl = impl_of(descA);
    auto* dimpl = impl_of(descdet);
    if (!dimpl->host) throw std::invalid_argument("descdet must be a host scalar descriptor");
    TileMatrix& tm = need_tm(impl);
    hipStream_t s = seq_stream(sequence);
    if (sequence && sequence->impl) {
      auto* simpl = static_cast<SeqImpl*>(sequence->impl);
      // Per-sequence partial buffer lives with the pending entry's host copy.
      double* dpart = rt().partial(tm.mt);
      launch_logdet(tm.view(), dpart, s);
      SeqImpl::Pending p;
      p.h.resize(tm.mt);
      p.target = dimpl->host;
      EXA_HIP(hipMemcpyAsync(p.h.data(), dpart, sizeof(double) * tm.mt, hipMemcpyDeviceToHost, s));
      EXA_HIP(hipStreamSynchronize(s));
      simpl->pending.push_back(std::move(p));
    } else {
      *dimpl->host += tile_logdet(tm, rt().partial(tm.mt), s);
    }
    return 0;
  }, -1);
}

int MORSE_MLE_dmdet_Tile_Async(CHAM_desc_t* descA, RUNTIME_sequence_t* sequence, RUNTIME_request_t* request,
                               CHAM_desc_t* descdet) {
  return EXAGEOSTAT_MLE_dmdet_Tile_Async(descA, sequence, request, descdet);
}

// ------------------------------------------------------------ MLE driver --
void EXAGEOSTAT_dmle_Call(MLE_data* data, int ncores, int gpus, int dts, int p_grid, int q_grid, int N,
                          int nZobs, int nZmiss) {
  (void)ncores;
  (void)gpus;
  (void)p_grid;
  (void)q_grid;
  guarded([&]() {
    if (data->gpu) MLE_Finalize(data);
    auto* ctx = new Context(N, dts, nZobs, nZmiss, data->mp_lean != 0);
    data->gpu = ctx;
    data->descC = make_desc(&ctx->C, false);
    data->descZ = make_vec_desc(&ctx->Z, dts);
    data->descZcpy = make_vec_desc(&ctx->Zcpy, dts);
    data->det = 0.0;
    data->dotp = 0.0;
    data->descdet = make_user_desc(&data->det, 1, 1);
    data->descproduct = make_user_desc(&data->dotp, 1, 1);
    data->iter_count = 0;
    data->status = 0;
    auto* seq = new RUNTIME_sequence_t();
    auto* simpl = new SeqImpl();
    simpl->s = ctx->s1;
    simpl->owns = false;
    seq->impl = simpl;
    seq->status = 0;
    data->sequence = seq;
    data->request = new RUNTIME_request_t();
    return 0;
  }, 0);
}

// The single-precision, mixed-precision and band-limited paths share the
// device context of EXAGEOSTAT_dmle_Call: the tile precision of each
// evaluation is chosen when it runs.
void EXAGEOSTAT_smle_Call(MLE_data* data, int ncores, int gpus, int dts, int p_grid, int q_grid, int N, int nZobs,
                          int nZmiss) {
  EXAGEOSTAT_dmle_Call(data, ncores, gpus, dts, p_grid, q_grid, N, nZobs, nZmiss);
}

void EXAGEOSTAT_sdmle_Call(MLE_data* data, int ncores, int gpus, int dts, int p_grid, int q_grid, int N, int nZobs,
                           int nZmiss) {
  EXAGEOSTAT_dmle_Call(data, ncores, gpus, dts, p_grid, q_grid, N, nZobs, nZmiss);
}

void EXAGEOSTAT_dmle_diag_Call(MLE_data* data, int ncores, int gpus, int dts, int p_grid, int q_grid, int N,
                               int nZobs, int nZmiss) {
  EXAGEOSTAT_dmle_Call(data, ncores, gpus, dts, p_grid, q_grid, N, nZobs, nZmiss);
}

void MLE_Finalize(MLE_data* data) {
  guarded([&]() {
    if (!data) return 0;
    release_prediction_descs(data);
    free_desc(data->descC);
    free_desc(data->descZ);
    free_desc(data->descZcpy);
    free_desc(data->descdet);
    free_desc(data->descproduct);
    if (data->sequence) {
      EXAGEOSTAT_Sequence_Destroy(static_cast<RUNTIME_sequence_t*>(data->sequence));
      data->sequence = nullptr;
    }
    delete static_cast<RUNTIME_request_t*>(data->request);
    data->request = nullptr;
    delete static_cast<Context*>(data->gpu);
    data->gpu = nullptr;
    return 0;
  }, 0);
}

void MLE_get_zobs(MLE_data* data, double* z, int n) {
  guarded([&]() {
    Context* ctx = ctx_of(data);
    EXA_HIP(hipMemcpyAsync(z, ctx->Zcpy.d, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->s1));
    EXA_HIP(hipStreamSynchronize(ctx->s1));
    return 0;
  }, 0);
}

location* GenerateXYLoc(int n, int seed) {
  srand(seed);
  auto* loc = static_cast<location*>(malloc(sizeof(location)));
  loc->x = static_cast<double*>(malloc(sizeof(double) * n));
  loc->y = static_cast<double*>(malloc(sizeof(double) * n));
  loc->z = nullptr;
  const int sqrtn = (int)ceil(sqrt((double)n));
  int index = 0;
  for (int i = 0; i < sqrtn && index < n; ++i)
    for (int j = 0; j < sqrtn && index < n; ++j) {
      // uniform_distribution(-0.4, 0.4), x drawn before y.
      const double ux = (double)rand() / (1.0 + RAND_MAX) * 0.8 - 0.4;
      loc->x[index] = ((i + 1) - 0.5 + ux) / sqrtn;
      const double uy = (double)rand() / (1.0 + RAND_MAX) * 0.8 - 0.4;
      loc->y[index] = ((j + 1) - 0.5 + uy) / sqrtn;
      ++index;
    }
  // Morton sort with 16-bit quantisation (zsort_locations).
  std::vector<uint32_t> code(n);
  for (int i = 0; i < n; ++i) {
    const uint16_t qx = (uint16_t)(loc->x[i] * 65535.0 + 0.5);
    const uint16_t qy = (uint16_t)(loc->y[i] * 65535.0 + 0.5);
    code[i] = (part1by1(qy) << 1) + part1by1(qx);
  }
  std::sort(code.begin(), code.end());
  for (int i = 0; i < n; ++i) {
    loc->x[i] = (double)compact1by1(code[i]) / 65535.0;
    loc->y[i] = (double)compact1by1(code[i] >> 1) / 65535.0;
  }
  return loc;
}

location* GenerateXYLoc_ST(int n, int t_slots, int seed) {
  srand(seed);
  const long total = (long)n * (t_slots > 0 ? t_slots : 0);
  auto* loc = static_cast<location*>(malloc(sizeof(location)));
  loc->x = static_cast<double*>(malloc(sizeof(double) * (total > 0 ? total : 1)));
  loc->y = static_cast<double*>(malloc(sizeof(double) * (total > 0 ? total : 1)));
  loc->z = static_cast<double*>(malloc(sizeof(double) * (total > 0 ? total : 1)));
  if (t_slots <= 0) return loc;
  const int sqrtn = (int)ceil(sqrt((double)n));
  int index = 0;
  for (int i = 0; i < sqrtn && index < n; ++i)
    for (int j = 0; j < sqrtn && index < n; ++j) {
      const double ux = (double)rand() / (1.0 + RAND_MAX) * 0.8 - 0.4;
      loc->x[index] = ((i + 1) - 0.5 + ux) / sqrtn;
      const double uy = (double)rand() / (1.0 + RAND_MAX) * 0.8 - 0.4;
      loc->y[index] = ((j + 1) - 0.5 + uy) / sqrtn;
      loc->z[index] = 1.0;
      ++index;
    }
  // No Morton sort: the spatial points of slot 0 repeat in every slot, with
  // time slot + 1.
  for (int t = 1; t < t_slots; ++t)
    for (int i = 0; i < n; ++i) {
      loc->x[i + (long)t * n] = loc->x[i];
      loc->y[i + (long)t * n] = loc->y[i];
      loc->z[i + (long)t * n] = (double)(t + 1);
    }
  return loc;
}

void pick_random_points(MLE_data* data, double* Zobs, double* Zactual, int nZmiss, int nZobs, int N) {
  guarded([&]() {
    std::vector<double> Z(N), x(N), y(N);
    MLE_get_zobs(data, Z.data(), N);
    for (int i = 0; i < N; ++i) {
      x[i] = data->l1.x[i];
      y[i] = data->l1.y[i];
    }
    // shuffle (misc/compute/MLE_misc.c:1238-1259), same rand() stream.
    if (N > 1) {
      for (size_t i = 0; i < (size_t)N - 1; ++i) {
        const size_t j = i + rand() / (RAND_MAX / ((size_t)N - i) + 1);
        std::swap(Z[i], Z[j]);
        std::swap(x[i], x[j]);
        std::swap(y[i], y[j]);
      }
    }
    for (int i = 0; i < nZobs; ++i) Zobs[i] = Z[nZmiss + i];
    for (int i = 0; i < nZmiss; ++i) Zactual[i] = Z[i];
    auto* mx = static_cast<double*>(malloc(sizeof(double) * (nZmiss > 0 ? nZmiss : 1)));
    auto* my = static_cast<double*>(malloc(sizeof(double) * (nZmiss > 0 ? nZmiss : 1)));
    auto* ox = static_cast<double*>(malloc(sizeof(double) * (nZobs > 0 ? nZobs : 1)));
    auto* oy = static_cast<double*>(malloc(sizeof(double) * (nZobs > 0 ? nZobs : 1)));
    std::copy(x.begin(), x.begin() + nZmiss, mx);
    std::copy(y.begin(), y.begin() + nZmiss, my);
    std::copy(x.begin() + nZmiss, x.begin() + nZmiss + nZobs, ox);
    std::copy(y.begin() + nZmiss, y.begin() + nZmiss + nZobs, oy);
    obs_zsort(nZobs, ox, oy, Zobs);
    obs_zsort(nZmiss, mx, my, Zactual);
    data->lmiss.x = mx;
    data->lmiss.y = my;
    data->lmiss.z = nullptr;
    data->lobs.x = ox;
    data->lobs.y = oy;
    data->lobs.z = nullptr;
    return 0;
  }, 0);
}

void EXAGEOSTAT_MLE_dzvg_Tile(MLE_data* data, double* Nrand, double* initial_theta, int n, int dts, int log) {
  (void)dts;
  (void)log;
  guarded([&]() {
    Context* ctx = ctx_of(data);
    const KernelId kid = kernel_from_name(data->kernel_fun);
    if (kid == kUnknownKernel) throw std::invalid_argument("unsupported kernel");
    const long nloc = n / kernel_vars(kid);
    ctx->l1.upload(data->l1.x, data->l1.y, nloc, ctx->s1, data->l1.z);
    TileMatrix& C = ctx->fp64("drawing z (MLE_zvg)");
    generate_covariance(C, kid, initial_theta, ctx->l1, ctx->l1, metric_of(data->dm), ctx->s1);
    const int info = tile_potrf(C, ctx->ws, ctx->s1, ctx->s2, true);
    if (info != 0) throw std::runtime_error("dzvg: covariance is not positive definite");
    copy_vec_to_device(ctx->work, Nrand, n, ctx->s1);
    ctx->gemv_part.allocate((long)ctx->C.mt * ctx->C.mt * ctx->C.ts, (long)ctx->C.mt * ctx->C.mt * ctx->C.ts);
    EXA_HIP(hipMemsetAsync(ctx->Z.d, 0, sizeof(double) * ctx->Z.padded, ctx->s1));
    launch_gemv_full(ctx->C.view(), ctx->work.d, ctx->Z.d, ctx->gemv_part.d, ctx->s1);
    EXA_HIP(hipMemcpyAsync(ctx->Zcpy.d, ctx->Z.d, sizeof(double) * ctx->Z.padded, hipMemcpyDeviceToDevice,
                           ctx->s1));
    EXA_HIP(hipStreamSynchronize(ctx->s1));
    ctx->gemv_part.release();
    data->iter_count = 0;
    return 0;
  }, 0);
}

void EXAGEOSTAT_MLE_dzvg_Tile_Async(MLE_data* data, double* Nrand, double* initial_theta, int n, int dts, int log) {
  EXAGEOSTAT_MLE_dzvg_Tile(data, Nrand, initial_theta, n, dts, log);
}

void MLE_zvg(MLE_data* data, double* Nrand, double* initial_theta, int n, int ts, int log, int p_grid,
             int q_grid) {
  (void)p_grid;
  (void)q_grid;
  data->iter_count = 0;
  EXAGEOSTAT_MLE_dzvg_Tile(data, Nrand, initial_theta, n, ts, log);
}

double EXAGEOSTAT_dmle_Tile(unsigned n, const double* theta, double* grad, void* vdata) {
  return evaluate(static_cast<MLE_data*>(vdata), n, theta, grad, Path::kExact);
}

double EXAGEOSTAT_dmle_Tile_Async(unsigned n, const double* theta, double* grad, void* vdata) {
  return evaluate(static_cast<MLE_data*>(vdata), n, theta, grad, Path::kExact);
}

double MORSE_dmle_Tile(unsigned n, const double* theta, double* grad, void* data) {
  return EXAGEOSTAT_dmle_Tile(n, theta, grad, data);
}

double EXAGEOSTAT_sdmle_Tile(unsigned n, const double* theta, double* grad, void* vdata) {
  return evaluate(static_cast<MLE_data*>(vdata), n, theta, grad, Path::kMixed);
}

double EXAGEOSTAT_sdmle_Tile_Async(unsigned n, const double* theta, double* grad, void* vdata) {
  return evaluate(static_cast<MLE_data*>(vdata), n, theta, grad, Path::kMixed);
}

double EXAGEOSTAT_smle_Tile(unsigned n, const double* theta, double* grad, void* vdata) {
  return evaluate(static_cast<MLE_data*>(vdata), n, theta, grad, Path::kSingle);
}

double EXAGEOSTAT_smle_Tile_Async(unsigned n, const double* theta, double* grad, void* vdata) {
  return evaluate(static_cast<MLE_data*>(vdata), n, theta, grad, Path::kSingle);
}

double EXAGEOSTAT_dmle_diag_Tile(unsigned n, const double* theta, double* grad, void* vdata) {
  return evaluate(static_cast<MLE_data*>(vdata), n, theta, grad, Path::kBand);
}

double EXAGEOSTAT_dmle_diag_Tile_Async(unsigned n, const double* theta, double* grad, void* vdata) {
  return evaluate(static_cast<MLE_data*>(vdata), n, theta, grad, Path::kBand);
}

double MLE_alg(unsigned n, const double* theta, double* grad, void* vdata) {
  auto* data = static_cast<MLE_data*>(vdata);
  const char* comp = data->computation ? data->computation : "exact";
  const bool exact = !strcmp(comp, "exact");
  if (exact && data->precision == 0) return EXAGEOSTAT_dmle_Tile(n, theta, grad, vdata);
  if (exact && data->precision == 1) return EXAGEOSTAT_smle_Tile(n, theta, grad, vdata);
  if (exact && data->precision == 2) return EXAGEOSTAT_sdmle_Tile(n, theta, grad, vdata);
  if (!strcmp(comp, "diag_approx") && data->precision == 0) return EXAGEOSTAT_dmle_diag_Tile(n, theta, grad, vdata);
  set_error(-1, std::string("unsupported combination: computation \"") + comp + "\" with precision " +
                    std::to_string(data->precision) +
                    " (exact: precision 0, 1 or 2; diag_approx: precision 0)");
  data->status = -1;
  return -HUGE_VAL;
}

// ------------------------------------------------------------- prediction --
void EXAGEOSTAT_dmle_Predict_Allocate(MLE_data* data, int nZmiss, int nZobs, int dts, int p_grid, int q_grid,
                                      int mse_flag) {
  (void)p_grid;
  (void)q_grid;
  (void)mse_flag;
  guarded([&]() {
    if (nZmiss <= 0) throw std::invalid_argument("Number of missing values should be positive value");
    Context* ctx = ctx_of(data);
    if (is_bivariate(data->kernel_fun)) {
      nZobs *= 2;
      nZmiss *= 2;
    }
    release_prediction_descs(data);
    ctx->C22.allocate(nZobs, nZobs, dts, true);
    ctx->C12.allocate(nZmiss, nZobs, dts, false);
    ctx->Zobs.allocate(nZobs, (long)ctx->C22.mt * dts);
    ctx->Zmiss.allocate(nZmiss, (long)ctx->C12.mt * dts);
    ctx->Zactual.allocate(nZmiss, (long)ctx->C12.mt * dts);
    ctx->gemv_part.allocate((long)ctx->C12.nt * ctx->C12.mt * dts, (long)ctx->C12.nt * ctx->C12.mt * dts);
    // The descriptors of MLE_dexact.c:655-683 over this context's buffers;
    // the MSE descriptors alias data->mserror{,1,2} as there.
    data->descC22 = make_desc(&ctx->C22, false);
    data->descC12 = make_desc(&ctx->C12, false);
    data->descZobs = make_vec_desc(&ctx->Zobs, dts);
    data->descZmiss = make_vec_desc(&ctx->Zmiss, dts);
    data->descZactual = make_vec_desc(&ctx->Zactual, dts);
    data->descmse = make_user_desc(&data->mserror, 1, 1);
    data->descmse1 = make_user_desc(&data->mserror1, 1, 1);
    data->descmse2 = make_user_desc(&data->mserror2, 1, 1);
    return 0;
  }, 0);
}

void prediction_init(MLE_data* data, int nZmiss, int nZobs, int ts, int p_grid, int q_grid, int mse_flag) {
  EXAGEOSTAT_dmle_Predict_Allocate(data, nZmiss, nZobs, ts, p_grid, q_grid, mse_flag);
}

void prediction_finalize(MLE_data* data) {
  guarded([&]() {
    if (!data || !data->gpu) return 0;
    Context* ctx = ctx_of(data);
    release_prediction_descs(data);
    ctx->C22.release();
    ctx->C12.release();
    ctx->Zobs.release();
    ctx->Zmiss.release();
    ctx->Zactual.release();
    ctx->gemv_part.release();
    return 0;
  }, 0);
}

// The prediction entry points of the FP64, single- and mixed-precision
// paths (mopt null: FP64).
static double predict_entry(MLE_data* data, double* theta, int nZmiss, int nZobs, double* Zobs, double* Zactual,
                            double* Zmiss, const MixedOptions* mopt) {
  if (nZmiss <= 0) {
    fprintf(stderr, " Number of missing values should be positive value\n");
    return -1;
  }
  return guarded([&]() {
    Context* ctx = ctx_of(data);
    const KernelId kid = kernel_from_name(data->kernel_fun);
    if (kid == kUnknownKernel) throw std::invalid_argument("unsupported kernel");
    const int vars = kernel_vars(kid);
    data->status = 0;
    predict_core(data, ctx, kid, theta, nZmiss, nZobs, data->lmiss, data->lobs, Zobs, mopt);
    hipStream_t s = ctx->s1;
    EXA_HIP(hipMemcpyAsync(Zmiss, ctx->Zmiss.d, sizeof(double) * nZmiss, hipMemcpyDeviceToHost, s));
    EXA_HIP(hipStreamSynchronize(s));
    if (!Zactual) {
      data->mserror = -1;
      return -1.0;
    }
    return prediction_mse(data, kid, vars, Zactual, Zmiss, nZmiss);
  }, -1.0);
}

double EXAGEOSTAT_dmle_Predict_Tile(MLE_data* data, double* theta, int nZmiss, int nZobs, double* Zobs,
                                    double* Zactual, double* Zmiss, int n) {
  (void)n;
  return predict_entry(data, theta, nZmiss, nZobs, Zobs, Zactual, Zmiss, nullptr);
}

double MORSE_dmle_Predict_Tile(MLE_data* data, double* theta, int nZmiss, int nZobs, double* Zobs,
                               double* Zactual, double* Zmiss, int n) {
  return EXAGEOSTAT_dmle_Predict_Tile(data, theta, nZmiss, nZobs, Zobs, Zactual, Zmiss, n);
}

// Single precision (MLE_sexact.c's prediction): C22 factored with every
// off-diagonal tile FP32 (diagonal tiles FP64), solved with that factor.
double EXAGEOSTAT_smle_Predict_Tile(MLE_data* data, double* theta, int nZmiss, int nZobs, double* Zobs,
                                    double* Zactual, double* Zmiss, int n) {
  (void)n;
  const MixedOptions opt = mixed_options_of(data, true);
  return predict_entry(data, theta, nZmiss, nZobs, Zobs, Zactual, Zmiss, &opt);
}

// Mixed precision (MLE_sdexact.c:528-680): C22 factored under the mp_*
// options of the mixed likelihood, the solve refined in FP64.
double EXAGEOSTAT_sdmle_Predict_Tile(MLE_data* data, double* theta, int nZmiss, int nZobs, double* Zobs,
                                     double* Zactual, double* Zmiss, int n) {
  (void)n;
  const MixedOptions opt = mixed_options_of(data, false);
  return predict_entry(data, theta, nZmiss, nZobs, Zobs, Zactual, Zmiss, &opt);
}

// The band-limited path's prediction (MLE_approx.c) is the FP64 kriging here.
double EXAGEOSTAT_dmle_diag_Predict_Tile(MLE_data* data, double* theta, int nZmiss, int nZobs, double* Zobs,
                                         double* Zactual, double* Zmiss, int n) {
  return EXAGEOSTAT_dmle_Predict_Tile(data, theta, nZmiss, nZobs, Zobs, Zactual, Zmiss, n);
}

// MLE_dexact.c:876-1000: the held-out locations are the first nZmiss of
// data->l1 and the observed ones the rest; Zobs and Zactual are read from
// data->descZobs / data->descZactual (filled by the caller, e.g. with
// EXAGEOSTAT_Lapack_to_Tile) and the predictions left in data->descZmiss.
// Returns the MSE, or -1 when there are no actual values.
double EXAGEOSTAT_dmle_Predict_Tile_Async(MLE_data* data, double* theta, int nZmiss, int nZobs, int n) {
  (void)n;
  if (nZmiss <= 0) {
    fprintf(stderr, " Number of missing values should be positive value\n");
    return -1;
  }
  return guarded([&]() {
    Context* ctx = ctx_of(data);
    const KernelId kid = kernel_from_name(data->kernel_fun);
    if (kid == kUnknownKernel) throw std::invalid_argument("unsupported kernel");
    if (!data->descZobs || !ctx->Zobs.d)
      throw std::invalid_argument("call EXAGEOSTAT_dmle_Predict_Allocate and fill descZobs first");
    const int vars = kernel_vars(kid);
    const long nmiss_loc = nZmiss / vars;
    location lmiss = data->l1;
    location lobs{data->l1.x + nmiss_loc, data->l1.y + nmiss_loc, data->l1.z ? data->l1.z + nmiss_loc : nullptr};
    predict_core(data, ctx, kid, theta, nZmiss, nZobs, lmiss, lobs, nullptr);
    if (!data->descZactual) {
      data->mserror = -1;
      return -1.0;
    }
    std::vector<double> zm(nZmiss), za(nZmiss);
    hipStream_t s = ctx->s1;
    EXA_HIP(hipMemcpyAsync(zm.data(), ctx->Zmiss.d, sizeof(double) * nZmiss, hipMemcpyDeviceToHost, s));
    EXA_HIP(hipMemcpyAsync(za.data(), ctx->Zactual.d, sizeof(double) * nZmiss, hipMemcpyDeviceToHost, s));
    EXA_HIP(hipStreamSynchronize(s));
    return prediction_mse(data, kid, vars, za.data(), zm.data(), nZmiss);
  }, -1.0);
}

// ------------------------------------------------------- synthetic data --
// LAPACK dlarnv: dlaruv's multiplicative congruential generator modulo 2^48
// with multiplier a = 33952834046453 (u_i = s a^i mod 2^48 / 2^48 for
// i = 1..m per call of at most 128 numbers), Box-Muller for idist = 3 in
// chunks of 64.
void exageostat_dlarnv(int idist, int* iseed, int n, double* x) {
  const uint64_t mask = (1ull << 48) - 1;
  const uint64_t a = 33952834046453ull;
  uint64_t mult[128];
  uint64_t p = 1;
  for (int i = 0; i < 128; ++i) {
    p = (uint64_t)(((unsigned __int128)p * a) & mask);
    mult[i] = p;
  }
  uint64_t s = ((uint64_t)iseed[0] << 36) | ((uint64_t)iseed[1] << 24) | ((uint64_t)iseed[2] << 12) |
               (uint64_t)iseed[3];
  double u[128];
  const double twopi = 6.28318530717958647692528676655900576839;
  for (int iv = 0; iv < n; iv += 64) {
    const int il = std::min(64, n - iv);
    const int il2 = idist == 3 ? 2 * il : il;
    for (int i = 0; i < il2; ++i)
      u[i] = (double)(uint64_t)(((unsigned __int128)s * mult[i]) & mask) / 281474976710656.0;
    s = (uint64_t)(((unsigned __int128)s * mult[il2 - 1]) & mask);
    for (int i = 0; i < il; ++i) {
      if (idist == 1) x[iv + i] = u[i];
      else if (idist == 2) x[iv + i] = 2.0 * u[i] - 1.0;
      else x[iv + i] = std::sqrt(-2.0 * std::log(u[2 * i])) * std::cos(twopi * u[2 * i + 1]);
    }
  }
  iseed[0] = (int)((s >> 36) & 4095);
  iseed[1] = (int)((s >> 24) & 4095);
  iseed[2] = (int)((s >> 12) & 4095);
  iseed[3] = (int)(s & 4095);
}

// ------------------------------------------------------------ measurement --
void exageostat_gpu_synchronize(void) {
  guarded([&]() {
    EXA_HIP(hipDeviceSynchronize());
    return 0;
  }, 0);
}

long exageostat_mixed_prec_map(MLE_data* data, unsigned char* out, long cap) {
  return guarded([&]() -> long {
    if (!data || !data->gpu) return 0;
    return mixed_prec_map(*static_cast<Context*>(data->gpu), out, cap);
  }, -1L);
}

int exageostat_fisher(MLE_data* data, const double* theta, double* fisher, double* fisher_inv) {
  return guarded([&]() {
    Context* ctx = ctx_of(data);
    const KernelId kid = kernel_from_name(data->kernel_fun);
    if (kid != kUnivariate && kid != kNugget)
      throw std::invalid_argument("Fisher information: univariate Matern kernels only");
    ctx->l1.upload(data->l1.x, data->l1.y, ctx->N, ctx->s1);
    const int np = kernel_num_params(kid);
    std::vector<double> F((size_t)np * np);
    const int info = fisher_information(*ctx, kid, theta, metric_of(data->dm), F.data());
    data->status = info;
    if (info != 0) {
      set_error(info, "covariance matrix is not positive definite");
      return info;
    }
    if (fisher) std::copy(F.begin(), F.end(), fisher);
    // F^-1 through the Cholesky factor of the small np x np matrix
    std::vector<double> L((size_t)np * np, 0.0), inv((size_t)np * np, 0.0);
    for (int j = 0; j < np; ++j) {
      double d = F[j * np + j];
      for (int k = 0; k < j; ++k) d -= L[j * np + k] * L[j * np + k];
      if (!(d > 0.0)) {
        if (fisher_inv) std::fill(fisher_inv, fisher_inv + np * np, std::nan(""));
        set_error(-2, "Fisher information matrix is not positive definite");
        return -2;
      }
      L[j * np + j] = std::sqrt(d);
      for (int i = j + 1; i < np; ++i) {
        double v = F[i * np + j];
        for (int k = 0; k < j; ++k) v -= L[i * np + k] * L[j * np + k];
        L[i * np + j] = v / L[j * np + j];
      }
    }
    for (int c = 0; c < np; ++c) {
      std::vector<double> y(np, 0.0);  // L y = e_c, then L^T x = y
      for (int i = 0; i < np; ++i) {
        double v = i == c ? 1.0 : 0.0;
        for (int k = 0; k < i; ++k) v -= L[i * np + k] * y[k];
        y[i] = v / L[i * np + i];
      }
      for (int i = np - 1; i >= 0; --i) {
        double v = y[i];
        for (int k = i + 1; k < np; ++k) v -= L[k * np + i] * inv[k * np + c];
        inv[i * np + c] = v / L[i * np + i];
      }
    }
    if (fisher_inv) std::copy(inv.begin(), inv.end(), fisher_inv);
    return 0;
  }, -1);
}

double* EXAGEOSTAT_Fisher_Tile(MLE_data* data, int N, double* initial_theta, int dts, int p_grid, int q_grid) {
  (void)dts;
  (void)p_grid;
  (void)q_grid;
  return guarded([&]() -> double* {
    Context* ctx = ctx_of(data);
    if (N != ctx->N) throw std::invalid_argument("Fisher: N differs from the context's size");
    const KernelId kid = kernel_from_name(data->kernel_fun);
    const int np = (kid == kUnivariate || kid == kNugget) ? kernel_num_params(kid) : 3;
    std::vector<double> inv((size_t)np * np);
    const double t0 = wall_seconds();
    if (exageostat_fisher(data, initial_theta, nullptr, inv.data()) != 0) return nullptr;
    const double t = wall_seconds() - t0;
    double* r = static_cast<double*>(calloc((size_t)np * np + 1, sizeof(double)));
    if (!r) throw std::runtime_error("Fisher: out of host memory");
    std::copy(inv.begin(), inv.end(), r);
    if (!data->silent) fprintf(stderr, " ---- Time to calculate Fisher matrix: %6.4f seconds\n", t);
    return r;
  }, static_cast<double*>(nullptr));
}

// ------------------------------------------------------------ MLOE / MMOM --
void EXAGEOSTAT_dmle_mloe_mmom_Allocate(MLE_data* data, int nZmiss, int nZobs, int dts, int p_grid, int q_grid) {
  (void)dts;
  (void)p_grid;
  (void)q_grid;
  guarded([&]() {
    if (nZmiss <= 0) throw std::invalid_argument("Number of missing values should be positive value");
    if (nZobs <= 0) throw std::invalid_argument("Number of observed values should be positive value");
    ctx_of(data);
    data->mloe = 0.0;
    data->mmom = 0.0;
    return 0;
  }, 0);
}

void mloe_mmom_init(MLE_data* data, int nZmiss, int nZobs, int ts, int p_grid, int q_grid) {
  EXAGEOSTAT_dmle_mloe_mmom_Allocate(data, nZmiss, nZobs, ts, p_grid, q_grid);
}

int exageostat_mloe_mmom(MLE_data* data, const double* truth_theta, const double* estimated_theta, int nZmiss,
                         int nZobs, double* loe, double* mom) {
  return guarded([&]() {
    Context* ctx = ctx_of(data);
    const KernelId kid = kernel_from_name(data->kernel_fun);
    if (kid == kUnknownKernel) throw std::invalid_argument("unsupported kernel");
    if (kid == kBivariate) throw std::invalid_argument("MLOE/MMOM: univariate Matern kernels only");
    if (nZmiss <= 0) throw std::invalid_argument("Number of missing values should be positive value");
    if (nZobs <= 0) throw std::invalid_argument("Number of observed values should be positive value");
    if (!data->lobs.x || !data->lobs.y || !data->lmiss.x || !data->lmiss.y)
      throw std::invalid_argument("MLOE/MMOM: data->lobs and data->lmiss must hold the locations");
    ctx->lobs.upload(data->lobs.x, data->lobs.y, nZobs, ctx->s1);
    ctx->lmiss.upload(data->lmiss.x, data->lmiss.y, nZmiss, ctx->s1);
    double mloe = 0.0, mmom = 0.0;
    const int info = mloe_mmom(*ctx, kid, truth_theta, estimated_theta, metric_of(data->dm), nZobs, nZmiss, loe,
                               mom, &mloe, &mmom);
    data->status = info;
    data->time_total = ctx->last.total;
    if (info != 0) {
      set_error(info, "MLOE/MMOM: covariance matrix is not positive definite");
      data->mloe = data->mmom = std::nan("");
      return info;
    }
    data->mloe = mloe;
    data->mmom = mmom;
    return 0;
  }, -1);
}

void EXAGEOSTAT_dmle_mloe_mmom_Tile(MLE_data* data, double* truth_theta, double* estimated_theta, int nZmiss,
                                    int nZobs, int n) {
  (void)n;
  const int rc = exageostat_mloe_mmom(data, truth_theta, estimated_theta, nZmiss, nZobs, nullptr, nullptr);
  if (rc < 0 && data) data->status = rc;
  if (rc == 0 && data && !data->silent) {
    printf("\nnZmiss = %d\n", nZmiss);
    printf("\nMMLOE = %f", data->mloe);
    printf("\nMMOM = %f\n\n\n", data->mmom);
    fprintf(stderr, " ---- mloe_mmom Time: %6.2f seconds\n\n", data->time_total);
  }
}

void EXAGEOSTAT_dmle_mloe_mmom_Tile_Async(MLE_data* data, double* truth_theta, double* estimated_theta, int nZmiss,
                                          int nZobs, int n) {
  EXAGEOSTAT_dmle_mloe_mmom_Tile(data, truth_theta, estimated_theta, nZmiss, nZobs, n);
}

void mloe_mmom_finalize(MLE_data* data) {
  guarded([&]() {
    if (!data || !data->gpu) return 0;
    Context* ctx = ctx_of(data);
    destroy_mloe(ctx->mloe);
    ctx->mloe = nullptr;
    return 0;
  }, 0);
}

void MLOE_MMOM_Finalize(MLE_data* data) { mloe_mmom_finalize(data); }

void exageostat_gpu_profile_enable(int on) { Profiler::get().enable(on != 0); }

void exageostat_gpu_profile_reset(void) { Profiler::get().reset(); }

int exageostat_gpu_profile_get(const char* name, long* launches, double* ms, double* flops, double* bytes) {
  return guarded([&]() {
    EXA_HIP(hipDeviceSynchronize());
    Profiler::get().collect();
    KernelTotals t = Profiler::get().totals(name);
    *launches = t.launches;
    *ms = t.ms;
    *flops = t.flops;
    *bytes = t.bytes;
    return 0;
  }, -1);
}

}  // extern "C"
