// Bound-constrained derivative-free maximisation by BOBYQA (M. J. D. Powell,
// "The BOBYQA algorithm for bound constrained optimization without
// derivatives", DAMTP 2009/NA06), with the calling conventions the reference
// uses through NLOPT: NLOPT_LN_BOBYQA as a max-objective
// (examples/synthetic_dmle_test.c:140,283), bounds and ftol_abs from
// init_optimizer (misc/compute/MLE_misc.c:1352-1366), maxeval from
// opt_max_iters, start at the lower bounds (examples/examples.c:263-264).
//
// The method as published: 2n+1 interpolation points, a quadratic model
// whose Hessian changes by the least Frobenius norm at each new point, a
// trust-region step under the bounds, a model-improvement ("alternative")
// step that moves a far interpolation point to where its Lagrange function
// is large, and a lower bound rho on the trust-region radius that is reduced
// when the model is good at the current scale.  Here the model and the
// Lagrange functions are obtained by solving the (m+n+1) x (m+n+1) KKT
// system of the least-change problem directly at every change of the point
// set (n is at most a few tens for these kernels), rather than by Powell's
// updating formulas of the inverse KKT matrix; the trust-region step is a
// truncated conjugate-gradient walk that fixes variables at the bounds it
// meets.  The initial step follows NLOPT's default: per variable the
// smallest of (ub - lb)/4, 3/4 of the distance to a bound, and the
// variables are rescaled so that every initial step equals the first.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/exageostat_gpu.h"

namespace {

using Vec = std::vector<double>;

double dot(const Vec& a, const Vec& b) {
  double s = 0.0;
  for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}

// Dense LU with partial pivoting of a k x k row-major matrix.
struct LU {
  int k = 0;
  Vec a;
  std::vector<int> piv;
  bool factor(int k_, const Vec& m) {
    k = k_;
    a = m;
    piv.resize(k);
    double scale = 0.0;
    for (double v : a) scale = std::max(scale, std::fabs(v));
    for (int c = 0; c < k; ++c) {
      int p = c;
      for (int r = c + 1; r < k; ++r)
        if (std::fabs(a[r * k + c]) > std::fabs(a[p * k + c])) p = r;
      piv[c] = p;
      if (!(std::fabs(a[p * k + c]) > 1e-14 * scale)) return false;
      if (p != c)
        for (int j = 0; j < k; ++j) std::swap(a[c * k + j], a[p * k + j]);
      for (int r = c + 1; r < k; ++r) {
        const double f = a[r * k + c] / a[c * k + c];
        a[r * k + c] = f;
        for (int j = c + 1; j < k; ++j) a[r * k + j] -= f * a[c * k + j];
      }
    }
    return true;
  }
  // Rows were swapped whole during the factorisation, so the permutation
  // is applied to b first, then the two triangular solves.
  void solve(Vec& b) const {
    for (int c = 0; c < k; ++c) std::swap(b[c], b[piv[c]]);
    for (int c = 0; c < k; ++c)
      for (int r = c + 1; r < k; ++r) b[r] -= a[r * k + c] * b[c];
    for (int c = k - 1; c >= 0; --c) {
      for (int j = c + 1; j < k; ++j) b[c] -= a[c * k + j] * b[j];
      b[c] /= a[c * k + c];
    }
  }
};

// A quadratic q(d) = c + g.d + d.H.d / 2 around the current best point.
struct Quad {
  double c = 0.0;
  Vec g;
  Vec H;  // n x n
  double eval(const Vec& d) const {
    const int n = (int)g.size();
    double v = c + dot(g, d);
    for (int i = 0; i < n; ++i) {
      double hd = 0.0;
      for (int j = 0; j < n; ++j) hd += H[i * n + j] * d[j];
      v += 0.5 * d[i] * hd;
    }
    return v;
  }
};

class Bobyqa {
 public:
  Bobyqa(int n, const Vec& sl, const Vec& su) : n_(n), m_(2 * n + 1), lo_(sl), hi_(su) {}

  // Interpolation points (absolute, scaled coordinates) and values of -f.
  std::vector<Vec> Y;
  Vec F;
  int kopt = 0;

  // Least-change model Q = Q_old + D: D interpolates F - Q_old at the points
  // and its Hessian sum_j lambda_j y_j y_j^T has the least Frobenius norm
  // (y_j relative to the best point).  Also keeps the factorised KKT matrix
  // for the Lagrange functions.  Returns false if the point set is
  // degenerate.
  bool rebuild(Quad& q, bool from_zero) {
    const int n = n_, m = m_;
    // shift the old model to the current best point
    const Vec& xo = Y[kopt];
    if (from_zero) {
      q.c = 0.0;
      q.g.assign(n, 0.0);
      q.H.assign((size_t)n * n, 0.0);
    } else {
      Vec s(n);
      for (int i = 0; i < n; ++i) s[i] = xo[i] - center_[i];
      const double cnew = q.eval(s);
      for (int i = 0; i < n; ++i) {
        double hs = 0.0;
        for (int j = 0; j < n; ++j) hs += q.H[i * n + j] * s[j];
        q.g[i] += hs;
      }
      q.c = cnew;
    }
    center_ = xo;
    ys_.assign(m, Vec(n, 0.0));
    sigma_ = 0.0;
    for (int j = 0; j < m; ++j) {
      for (int i = 0; i < n; ++i) ys_[j][i] = Y[j][i] - xo[i];
      sigma_ = std::max(sigma_, std::sqrt(dot(ys_[j], ys_[j])));
    }
    if (!(sigma_ > 0.0)) return false;
    for (auto& y : ys_)
      for (double& v : y) v /= sigma_;
    const int k = m + n + 1;
    Vec K((size_t)k * k, 0.0);
    for (int a = 0; a < m; ++a) {
      for (int b = 0; b < m; ++b) {
        const double t = dot(ys_[a], ys_[b]);
        K[a * k + b] = 0.5 * t * t;
      }
      K[a * k + m] = K[m * k + a] = 1.0;
      for (int i = 0; i < n; ++i) K[a * k + m + 1 + i] = K[(m + 1 + i) * k + a] = ys_[a][i];
    }
    if (!lu_.factor(k, K)) return false;
    Vec r(k, 0.0);
    for (int a = 0; a < m; ++a) {
      Vec d(n);
      for (int i = 0; i < n; ++i) d[i] = Y[a][i] - xo[i];
      r[a] = F[a] - q.eval(d);
    }
    lu_.solve(r);
    add_solution(q, r, 1.0);
    return true;
  }

  // Lagrange function t of the current point set, around the best point.
  Quad lagrange(int t) const {
    const int k = m_ + n_ + 1;
    Vec r(k, 0.0);
    r[t] = 1.0;
    lu_.solve(r);
    Quad q;
    q.c = 0.0;
    q.g.assign(n_, 0.0);
    q.H.assign((size_t)n_ * n_, 0.0);
    add_solution(q, r, 1.0);
    return q;
  }

  int n_, m_;
  Vec lo_, hi_;  // bounds (scaled)

 private:
  // Adds the quadratic of a KKT solution (lambda', c, g') in scaled
  // coordinates: Hessian sum lambda'_j yt_j yt_j^T / sigma^2, gradient
  // g' / sigma, constant c.
  void add_solution(Quad& q, const Vec& r, double w) const {
    const int n = n_, m = m_;
    q.c += w * r[m];
    for (int i = 0; i < n; ++i) q.g[i] += w * r[m + 1 + i] / sigma_;
    const double s2 = sigma_ * sigma_;
    for (int j = 0; j < m; ++j) {
      const double l = w * r[j] / s2;
      if (l == 0.0) continue;
      for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b) q.H[a * n + b] += l * ys_[j][a] * ys_[j][b];
    }
  }

  Vec center_;
  std::vector<Vec> ys_;
  double sigma_ = 1.0;
  LU lu_;
};

// Truncated conjugate gradients for min g.d + d.H.d/2 with ||d|| <= delta
// and sl <= d <= su (sl <= 0 <= su); a variable that reaches a bound is
// fixed there and the walk restarts on the others.
Vec trust_step(const Quad& q, double delta, const Vec& sl, const Vec& su) {
  const int n = (int)q.g.size();
  Vec d(n, 0.0), gd = q.g;
  std::vector<char> fixed(n, 0);
  for (int i = 0; i < n; ++i)
    if ((sl[i] >= 0.0 && gd[i] >= 0.0) || (su[i] <= 0.0 && gd[i] <= 0.0)) fixed[i] = 1;
  auto hv = [&](const Vec& v) {
    Vec o(n, 0.0);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) o[i] += q.H[i * n + j] * v[j];
    return o;
  };
  const double g0 = std::sqrt(dot(q.g, q.g));
  for (int restart = 0; restart <= n; ++restart) {
    Vec r(n, 0.0);
    for (int i = 0; i < n; ++i) r[i] = fixed[i] ? 0.0 : -gd[i];
    double rr = dot(r, r);
    if (rr <= 1e-24 * g0 * g0) break;
    Vec s = r;
    bool restart_needed = false;
    for (int it = 0; it < n; ++it) {
      const Vec Hs = hv(s);
      const double sHs = dot(s, Hs);
      // step to the trust-region boundary
      const double ss = dot(s, s), ds = dot(d, s), dd = dot(d, d);
      const double disc = ds * ds + ss * std::max(0.0, delta * delta - dd);
      const double a_tr = (std::sqrt(disc) - ds) / ss;
      // step to the first bound
      double a_b = HUGE_VAL;
      int ib = -1;
      for (int i = 0; i < n; ++i) {
        if (fixed[i] || s[i] == 0.0) continue;
        const double lim = s[i] > 0.0 ? (su[i] - d[i]) / s[i] : (sl[i] - d[i]) / s[i];
        if (lim < a_b) {
          a_b = std::max(0.0, lim);
          ib = i;
        }
      }
      const double a_cg = sHs > 0.0 ? rr / sHs : HUGE_VAL;
      const double a = std::min(a_cg, std::min(a_tr, a_b));
      for (int i = 0; i < n; ++i) {
        d[i] += a * s[i];
        gd[i] += a * Hs[i];
      }
      if (a == a_b && ib >= 0) {
        d[ib] = s[ib] > 0.0 ? su[ib] : sl[ib];
        fixed[ib] = 1;
        restart_needed = true;
        break;
      }
      if (a == a_tr) return d;
      Vec rn(n, 0.0);
      for (int i = 0; i < n; ++i) rn[i] = fixed[i] ? 0.0 : -gd[i];
      const double rrn = dot(rn, rn);
      if (rrn <= 1e-4 * rr && it > 0) return d;
      const double beta = rrn / rr;
      for (int i = 0; i < n; ++i) s[i] = rn[i] + beta * s[i];
      rr = rrn;
    }
    if (!restart_needed) break;
    if (std::sqrt(dot(d, d)) >= delta * (1.0 - 1e-12)) break;
  }
  return d;
}

// Model-improvement step for interpolation point t: a step from the best
// point, of length at most adelt and inside the bounds, at which |l_t| is
// large; candidates along the lines to the other points and along +-grad l_t.
Vec alternative_step(const Bobyqa& B, const Quad& lt, double adelt, const Vec& sl, const Vec& su) {
  const int n = B.n_;
  Vec best(n, 0.0);
  double bestv = -1.0;
  auto consider = [&](const Vec& dir, double tmax_len) {
    const double len = std::sqrt(dot(dir, dir));
    if (!(len > 0.0)) return;
    double tlo = -tmax_len / len, thi = tmax_len / len;
    for (int i = 0; i < n; ++i) {
      if (dir[i] > 0.0) {
        thi = std::min(thi, su[i] / dir[i]);
        tlo = std::max(tlo, sl[i] / dir[i]);
      } else if (dir[i] < 0.0) {
        thi = std::min(thi, sl[i] / dir[i]);
        tlo = std::max(tlo, su[i] / dir[i]);
      }
    }
    if (!(thi >= tlo)) return;
    // l_t(t dir) = l0 + t gv + t^2 hv / 2
    double gv = dot(lt.g, dir), hvv = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) hvv += dir[i] * lt.H[i * n + j] * dir[j];
    std::vector<double> ts = {tlo, thi};
    if (hvv != 0.0) {
      const double tc = -gv / hvv;
      if (tc > tlo && tc < thi) ts.push_back(tc);
    }
    for (double t : ts) {
      const double v = std::fabs(lt.c + t * gv + 0.5 * t * t * hvv);
      if (v > bestv && t != 0.0) {
        bestv = v;
        for (int i = 0; i < n; ++i) best[i] = t * dir[i];
      }
    }
  };
  for (int j = 0; j < B.m_; ++j) {
    if (j == B.kopt) continue;
    Vec dir(n);
    for (int i = 0; i < n; ++i) dir[i] = B.Y[j][i] - B.Y[B.kopt][i];
    consider(dir, adelt);
  }
  consider(lt.g, adelt);
  return best;
}

}  // namespace

extern "C" int exageostat_bobyqa(unsigned nvar, exageostat_func_t* f, void* f_data, const double* lb, const double* ub,
                                 double* x, double ftol_abs, int maxeval, double* fmax, int* nevals) {
  if (!f || !lb || !ub || !x || nvar == 0) return -1;
  const int N = (int)nvar;
  for (int i = 0; i < N; ++i)
    if (!(lb[i] <= ub[i])) return -1;
  // variables with lb == ub are fixed (NLOPT eliminates them)
  std::vector<int> freev;
  for (int i = 0; i < N; ++i) {
    x[i] = std::min(ub[i], std::max(lb[i], x[i]));
    if (lb[i] < ub[i]) freev.push_back(i);
  }
  Vec xfull(x, x + N);
  int nf = 0;
  double best_f = -HUGE_VAL;
  Vec best_x = xfull;
  auto call = [&](const Vec& xf) {
    const double v = f(nvar, xf.data(), nullptr, f_data);
    ++nf;
    if (v > best_f || nf == 1) {
      best_f = v;
      best_x = xf;
    }
    return v;
  };
  auto finish = [&](int rc) {
    std::copy(best_x.begin(), best_x.end(), x);
    if (fmax) *fmax = best_f;
    if (nevals) *nevals = nf;
    return rc;
  };
  const int n = (int)freev.size();
  if (n == 0) {
    call(xfull);
    return finish(0);
  }
  // NLOPT's default initial step per variable, then rescaling so that all
  // initial steps equal the first one (rhobeg).
  Vec step(n);
  for (int k = 0; k < n; ++k) {
    const int i = freev[k];
    double s = 0.25 * (ub[i] - lb[i]);
    if (ub[i] - x[i] < s && ub[i] > x[i]) s = 0.75 * (ub[i] - x[i]);
    if (x[i] - lb[i] < s && x[i] > lb[i]) s = 0.75 * (x[i] - lb[i]);
    step[k] = s;
  }
  const double rhobeg = step[0];
  Vec scale(n);
  for (int k = 0; k < n; ++k) scale[k] = step[k] / rhobeg;
  auto to_full = [&](const Vec& u) {
    Vec xf = xfull;
    for (int k = 0; k < n; ++k) {
      const int i = freev[k];
      xf[i] = std::min(ub[i], std::max(lb[i], u[k] * scale[k]));
    }
    return xf;
  };
  Vec ulo(n), uhi(n), u0(n);
  for (int k = 0; k < n; ++k) {
    const int i = freev[k];
    ulo[k] = lb[i] / scale[k];
    uhi[k] = ub[i] / scale[k];
    u0[k] = x[i] / scale[k];
  }
  // Powell's start: a variable within rhobeg of a bound moves onto the bound
  // or rhobeg inside it, so that the initial points are feasible.
  for (int k = 0; k < n; ++k) {
    if (u0[k] - ulo[k] < rhobeg) u0[k] = (u0[k] - ulo[k] <= 0.0) ? ulo[k] : std::min(uhi[k], ulo[k] + rhobeg);
    else if (uhi[k] - u0[k] < rhobeg) u0[k] = (uhi[k] - u0[k] <= 0.0) ? uhi[k] : std::max(ulo[k], uhi[k] - rhobeg);
  }
  const bool limit = maxeval > 0;
  Bobyqa B(n, ulo, uhi);
  B.Y.assign(2 * n + 1, u0);
  B.F.assign(2 * n + 1, 0.0);
  // initial points: x0, x0 + a_k e_k, x0 + b_k e_k (PRELIM)
  for (int k = 0; k < n; ++k) {
    const double sl = ulo[k] - u0[k], su = uhi[k] - u0[k];
    double a = rhobeg;
    if (su <= 0.0) a = -rhobeg;
    double b = -rhobeg;
    if (sl >= 0.0) b = std::min(2.0 * rhobeg, su);
    if (su <= 0.0) b = std::max(-2.0 * rhobeg, sl);
    if (a == b) b = -a;
    B.Y[1 + k][k] += a;
    B.Y[1 + n + k][k] += b;
  }
  for (int j = 0; j < 2 * n + 1; ++j) {
    B.F[j] = -call(to_full(B.Y[j]));
    if (!std::isfinite(B.F[j])) B.F[j] = HUGE_VAL;
    if (limit && nf >= maxeval) return finish(1);
  }
  B.kopt = (int)(std::min_element(B.F.begin(), B.F.end()) - B.F.begin());
  Quad q;
  if (!B.rebuild(q, true)) return finish(-2);
  double rho = rhobeg, delta = rhobeg;
  const double rho_floor = 1e-14 * rhobeg;
  auto bounds_rel = [&](Vec& sl, Vec& su) {
    sl.resize(n);
    su.resize(n);
    for (int k = 0; k < n; ++k) {
      sl[k] = ulo[k] - B.Y[B.kopt][k];
      su[k] = uhi[k] - B.Y[B.kopt][k];
    }
  };
  // replace point t by u with value fu (minimisation of -f) and refresh
  auto replace = [&](int t, const Vec& u, double fu) {
    const double fold = B.F[B.kopt];
    B.Y[t] = u;
    B.F[t] = fu;
    if (fu < fold) B.kopt = t;
    else if (t == B.kopt) B.kopt = (int)(std::min_element(B.F.begin(), B.F.end()) - B.F.begin());
    return B.rebuild(q, false);
  };
  auto farthest = [&](double& d2max) {
    int t = -1;
    d2max = 0.0;
    for (int j = 0; j < B.m_; ++j) {
      double d2 = 0.0;
      for (int k = 0; k < n; ++k) d2 += (B.Y[j][k] - B.Y[B.kopt][k]) * (B.Y[j][k] - B.Y[B.kopt][k]);
      if (d2 > d2max) {
        d2max = d2;
        t = j;
      }
    }
    return t;
  };
  for (;;) {
    Vec sl, su;
    bounds_rel(sl, su);
    const Vec d = trust_step(q, delta, sl, su);
    const double dnorm = std::sqrt(dot(d, d));
    bool try_geometry = false;
    double geo_dist2 = 0.0;
    if (dnorm < 0.5 * rho) {
      // the model predicts too small a change: improve the geometry at
      // this scale, else reduce rho
      delta = std::max(rho, 0.5 * delta);
      if (delta <= 1.5 * rho) delta = rho;
      try_geometry = true;
      geo_dist2 = 100.0 * rho * rho;
    } else {
      Vec un(n);
      for (int k = 0; k < n; ++k) un[k] = std::min(uhi[k], std::max(ulo[k], B.Y[B.kopt][k] + d[k]));
      const double fopt = B.F[B.kopt];
      double fn = -call(to_full(un));
      if (!std::isfinite(fn)) fn = HUGE_VAL;
      // NLOPT's ftol_abs test: the new value within ftol_abs of the best
      if (std::fabs(fn - fopt) < ftol_abs) return finish(0);
      if (limit && nf >= maxeval) return finish(1);
      const double pred = -q.eval(d) + q.c;  // predicted decrease Q(0) - Q(d)
      const double ratio = pred > 0.0 ? (fopt - fn) / pred : -1.0;
      if (ratio <= 0.1) delta = std::min(0.5 * delta, dnorm);
      else if (ratio <= 0.7) delta = std::max(0.5 * delta, dnorm);
      else delta = std::max(0.5 * delta, 2.0 * dnorm);
      if (delta <= 1.5 * rho) delta = rho;
      // the point to drop: largest |l_t(new)| weighted by distance
      int knew = -1;
      double score = 0.0;
      for (int t = 0; t < B.m_; ++t) {
        if (t == B.kopt && fn >= fopt) continue;
        const Quad lt = B.lagrange(t);
        const double lv = std::fabs(lt.eval(d));
        double d2 = 0.0;
        for (int k = 0; k < n; ++k) d2 += (B.Y[t][k] - B.Y[B.kopt][k]) * (B.Y[t][k] - B.Y[B.kopt][k]);
        const double w = std::max(1.0, d2 / (delta * delta));
        const double sc = lv * w * w;
        if (sc > score) {
          score = sc;
          knew = t;
        }
      }
      if (knew >= 0 && score > 1e-12) {
        const std::vector<Vec> Ysave = B.Y;
        const Vec Fsave = B.F;
        const int ksave = B.kopt;
        const Quad qsave = q;
        if (!replace(knew, un, fn)) {
          B.Y = Ysave;
          B.F = Fsave;
          B.kopt = ksave;
          q = qsave;
          B.rebuild(q, false);
        }
      }
      if (ratio >= 0.1) continue;
      try_geometry = true;
      geo_dist2 = std::max(delta * delta, 4.0 * rho * rho);
    }
    if (try_geometry) {
      double d2max = 0.0;
      const int t = farthest(d2max);
      if (t >= 0 && d2max > geo_dist2) {
        const double adelt = std::max(std::min(0.1 * std::sqrt(d2max), delta), rho);
        Vec gsl, gsu;
        bounds_rel(gsl, gsu);
        const Quad lt = B.lagrange(t);
        const Vec dg = alternative_step(B, lt, adelt, gsl, gsu);
        if (std::sqrt(dot(dg, dg)) > 1e-3 * adelt) {
          Vec un(n);
          for (int k = 0; k < n; ++k) un[k] = std::min(uhi[k], std::max(ulo[k], B.Y[B.kopt][k] + dg[k]));
          const double fopt = B.F[B.kopt];
          double fn = -call(to_full(un));
          if (!std::isfinite(fn)) fn = HUGE_VAL;
          if (std::fabs(fn - fopt) < ftol_abs) return finish(0);
          if (limit && nf >= maxeval) return finish(1);
          const std::vector<Vec> Ysave = B.Y;
          const Vec Fsave = B.F;
          const int ksave = B.kopt;
          const Quad qsave = q;
          if (!replace(t, un, fn)) {
            B.Y = Ysave;
            B.F = Fsave;
            B.kopt = ksave;
            q = qsave;
            B.rebuild(q, false);
          } else {
            continue;
          }
        }
      }
      // the model is adequate at this scale: reduce rho
      if (rho <= rho_floor) return finish(2);
      const double rho_new = 0.1 * rho;
      delta = std::max(0.5 * rho, rho_new);
      rho = rho_new;
    }
  }
}
