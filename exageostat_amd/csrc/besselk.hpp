// FP64 modified Bessel function of the second kind K_nu(x), real nu >= 0,
// x > 0, for host and device, plus a forward-mode dual number so the same
// code yields dK_nu/dnu.
//
// Replaces GSL 2.6 gsl_sf_bessel_Knu, which the reference calls for every
// covariance entry (exageostat_exact/cpu_core/compute/core_zcmg.c:987).
//
// Method (published algorithms, own implementation):
//   nu = N + mu with N = floor(nu + 1/2), mu in [-1/2, 1/2).
//   x <= 2: Temme's series for K_mu and K_{mu+1}
//           (N. M. Temme, J. Comput. Phys. 19 (1975) 324-337).
//   x >  2: Steed's evaluation of the continued fraction CF2 with Temme's
//           normalisation (I. J. Thompson, A. R. Barnett, Comput. Phys.
//           Commun. 47 (1987) 245-257).
//   then the forward recurrence K_{a+1} = K_{a-1} + (2a/x) K_a up to nu+1.
// Everything that depends on nu only (Gamma_1(mu), Gamma_2(mu),
// 1/Gamma(1+-mu), mu*pi/sin(mu*pi) and the reciprocal tables of the series
// recurrences) is computed once per nu on the host (knu_prepare), so the
// per-entry device code has no divisions in the Temme branch.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace exa {

#define EXA_HD __host__ __device__ __forceinline__

// 1/y for the finite, normal, moderate y of the loops below.  Device: the
// hardware reciprocal estimate refined by two Newton steps with fused
// multiply-adds (within about one unit in the last place; no scaling or
// special-case fix-up, which these operands never need); host: the
// division.
EXA_HD double recip(double y) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(y);
  double e = ::fma(-y, r, 1.0);
  r = ::fma(r, e, r);
  e = ::fma(-y, r, 1.0);
  return ::fma(r, e, r);
#else
  return 1.0 / y;
#endif
}

// ---------------------------------------------------------------- dual --
struct Dual {
  double v, d;
  EXA_HD Dual() : v(0.0), d(0.0) {}
  EXA_HD Dual(double a) : v(a), d(0.0) {}
  EXA_HD Dual(double a, double b) : v(a), d(b) {}
};
EXA_HD Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
EXA_HD Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
EXA_HD Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
EXA_HD Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
EXA_HD Dual operator/(Dual a, Dual b) {
  const double r = recip(b.v);
  const double q = a.v * r;
  return Dual(q, (a.d - q * b.d) * r);
}
EXA_HD Dual operator+(Dual a, double b) { return Dual(a.v + b, a.d); }
EXA_HD Dual operator+(double a, Dual b) { return Dual(a + b.v, b.d); }
EXA_HD Dual operator-(Dual a, double b) { return Dual(a.v - b, a.d); }
EXA_HD Dual operator-(double a, Dual b) { return Dual(a - b.v, -b.d); }
EXA_HD Dual operator*(Dual a, double b) { return Dual(a.v * b, a.d * b); }
EXA_HD Dual operator*(double a, Dual b) { return Dual(a * b.v, a * b.d); }
EXA_HD Dual operator/(Dual a, double b) { return Dual(a.v / b, a.d / b); }
EXA_HD Dual operator/(double a, Dual b) {
  const double r = recip(b.v);
  const double q = a * r;
  return Dual(q, -q * b.d * r);
}
EXA_HD Dual& operator+=(Dual& a, Dual b) { a.v += b.v; a.d += b.d; return a; }
EXA_HD Dual& operator*=(Dual& a, double b) { a.v *= b; a.d *= b; return a; }
EXA_HD Dual exp(Dual a) { const double e = ::exp(a.v); return Dual(e, e * a.d); }
EXA_HD Dual log(Dual a) { return Dual(::log(a.v), a.d / a.v); }
EXA_HD Dual sqrt(Dual a) { const double s = ::sqrt(a.v); return Dual(s, 0.5 * a.d / s); }
EXA_HD double exp(double a) { return ::exp(a); }
EXA_HD double log(double a) { return ::log(a); }
EXA_HD double sqrt(double a) { return ::sqrt(a); }
// a / b and 1 / b through recip (the loop divisions of knu_steed)
EXA_HD double quot(double a, double b) { return a * recip(b); }
EXA_HD Dual quot(Dual a, Dual b) { return a / b; }
EXA_HD double inv(double b) { return recip(b); }
EXA_HD Dual inv(Dual b) { return 1.0 / b; }
EXA_HD double val(double a) { return a; }
EXA_HD double val(Dual a) { return a.v; }
EXA_HD double dot_d(Dual a) { return a.d; }
// Series/fraction convergence: a term is negligible when it is below eps
// relative to the sum; for a dual number the derivative part must be
// negligible too (the value part can vanish identically, e.g. CF2 at
// mu = -1/2, while its derivative does not).
EXA_HD bool negligible(double term, double sum, double eps) {
  return ::fabs(term) < ::fabs(sum) * eps;
}
EXA_HD bool negligible(Dual term, Dual sum, double eps) {
  return ::fabs(term.v) < ::fabs(sum.v) * eps &&
         ::fabs(term.d) < (::fabs(sum.d) + ::fabs(sum.v)) * eps;
}

// ----------------------------------------------------- per-order data --
constexpr int kTemmeTerms = 24;  // x <= 2 needs <= 14 terms for 1e-17
constexpr int kCf2MaxIter = 200;
constexpr double kKnuEps = 1.0e-16;

template <class T>
struct KnuOrder {
  T nu;             // the order
  T mu;             // nu - n_rec, in [-1/2, 1/2)
  T g1, g2;         // Temme's Gamma_1(mu), Gamma_2(mu)
  T gp, gm;         // Gamma(1+mu), Gamma(1-mu)
  T fact;           // mu*pi / sin(mu*pi)
  T a1;             // 1/4 - mu^2 (CF2)
  T rden[kTemmeTerms];  // 1/(i^2 - mu^2), i = 1..
  T rmm[kTemmeTerms];   // 1/(i - mu)
  T rmp[kTemmeTerms];   // 1/(i + mu)
  int n_rec;        // N in nu = N + mu
  int closed_form;  // 1: nu = 1/2, 3: nu = 3/2, 5: nu = 5/2, else 0
};

// ------------------------------------------------------- K_mu, K_mu+1 --
template <class T>
EXA_HD void knu_temme(const KnuOrder<T>& o, double x, T& kmu, T& kmu1) {
  const double x2 = 0.5 * x;
  const double lx = -::log(x2);  // ln(2/x) >= 0
  const T sig = o.mu * lx;
  const T e = exp(sig);            // (x/2)^(-mu)
  const T ei = 1.0 / e;
  const T ch = 0.5 * (e + ei);
  T shs;                           // sinh(sig)/sig
  if (::fabs(val(sig)) < 1.0e-3) {
    const T s2 = sig * sig;
    shs = 1.0 + s2 * (1.0 / 6.0 + s2 * (1.0 / 120.0 + s2 * (1.0 / 5040.0)));
  } else {
    shs = 0.5 * (e - ei) / sig;
  }
  T ff = o.fact * (o.g1 * ch + o.g2 * shs * lx);
  T p = 0.5 * e * o.gp;    // (1/2)(x/2)^-mu Gamma(1+mu)
  T q = 0.5 * ei * o.gm;   // (1/2)(x/2)^mu  Gamma(1-mu)
  const double dd = x2 * x2;
  double c = 1.0;
  T sum = ff;
  T sum1 = p;
#pragma unroll 1
  for (int i = 1; i <= kTemmeTerms; ++i) {
    ff = (double(i) * ff + p + q) * o.rden[i - 1];
    c *= dd * recip(double(i));
    p = p * o.rmm[i - 1];
    q = q * o.rmp[i - 1];
    const T del = c * ff;
    sum += del;
    const T del1 = c * (p - double(i) * ff);
    sum1 += del1;
    if (negligible(del, sum, kKnuEps)) break;
  }
  kmu = sum;
  kmu1 = sum1 * (2.0 / x);
}

template <class T>
EXA_HD void knu_steed(const KnuOrder<T>& o, double x, T& kmu, T& kmu1) {
  double b = 2.0 * (1.0 + x);
  T d = T(1.0 / b);
  T h = d;
  T delh = d;
  T q1 = T(0.0);
  T q2 = T(1.0);
  const T a1 = o.a1;
  T q = a1;
  T c = a1;
  T a = -a1;
  T s = 1.0 + q * delh;
#pragma unroll 1
  for (int i = 2; i <= kCf2MaxIter; ++i) {
    a = a - double(2 * (i - 1));
    c = -a * c * recip(double(i));
    const T qn = quot(q1 - b * q2, a);
    q1 = q2;
    q2 = qn;
    q += c * qn;
    b += 2.0;
    d = inv(b + a * d);
    delh = (b * d - 1.0) * delh;
    h += delh;
    const T dels = q * delh;
    s += dels;
    if (negligible(dels, s, kKnuEps)) break;
  }
  h = a1 * h;
  // sqrt(pi/(2x)) e^-x / s
  kmu = (1.2533141373155002512078826 / ::sqrt(x)) * ::exp(-x) / s;
  kmu1 = kmu * (o.mu + x + 0.5 - h) * (1.0 / x);
}

// Returns K_nu(x) and K_{nu+1}(x).
template <class T>
EXA_HD void knu_pair(const KnuOrder<T>& o, double x, T& knu, T& knu1) {
  T km, kp;
  if (x <= 2.0) {
    knu_temme(o, x, km, kp);
  } else {
    knu_steed(o, x, km, kp);
  }
  // Forward recurrence from (K_mu, K_mu+1) to (K_nu, K_nu+1).
  T a = o.mu + 1.0;
  const double rx = 2.0 / x;
#pragma unroll 1
  for (int k = 0; k < o.n_rec; ++k) {
    const T kn = km + a * rx * kp;
    km = kp;
    kp = kn;
    a = a + 1.0;
  }
  knu = km;
  knu1 = kp;
}

// ------------------------------------------------------------ host prep --
// Taylor coefficients of 1/Gamma(z) = sum_{k>=1} kRG[k] z^k, obtained with
// mpmath.taylor(mpmath.rgamma, 0, 32) at 40 digits.
constexpr int kRGTerms = 33;
constexpr double kRG[kRGTerms] = {
    0.0,
    1.0,
    0.5772156649015328606065121,
    -0.6558780715202538810770195,
    -0.04200263503409523552900393,
    0.1665386113822914895017008,
    -0.0421977345555443367482083,
    -0.009621971527876973562114922,
    0.00721894324666309954239501,
    -0.001165167591859065112113971,
    -0.00021524167411495097281573,
    0.0001280502823881161861531986,
    -0.00002013485478078823865568939,
    -0.000001250493482142670657345359,
    0.00000113302723198169588237413,
    -0.0000002056338416977607103450154,
    6.116095104481415817862499e-9,
    5.002007644469222930055665e-9,
    -1.181274570487020144588127e-9,
    1.04342671169110051049154e-10,
    7.782263439905071254049937e-12,
    -3.696805618642205708187816e-12,
    5.100370287454475979015481e-13,
    -2.05832605356650678322243e-14,
    -5.348122539423017982370017e-15,
    1.226778628238260790158894e-15,
    -1.181259301697458769513765e-16,
    1.186692254751600332579777e-18,
    1.412380655318031781555804e-18,
    -2.298745684435370206592479e-19,
    1.714406321927337433383963e-20,
    1.337351730493693114864781e-22,
    -2.054233551766672789325025e-22,
};

// 1/Gamma(1+z) = sum_{k>=0} kRG[k+1] z^k and its derivative, |z| <= 1/2.
inline Dual host_rgamma1p(Dual z) {
  double s = 0.0, ds = 0.0;
  for (int k = kRGTerms - 2; k >= 0; --k) {
    ds = ds * z.v + s;
    s = s * z.v + kRG[k + 1];
  }
  return Dual(s, ds * z.d);
}

// Temme's Gamma_1(mu) = (1/Gamma(1-mu) - 1/Gamma(1+mu)) / (2 mu)
//                     = -sum_{j>=0} kRG[2j+2] mu^(2j),
//         Gamma_2(mu) = (1/Gamma(1-mu) + 1/Gamma(1+mu)) / 2
//                     =  sum_{j>=0} kRG[2j+1] mu^(2j).
inline void host_temme_gammas(Dual mu, Dual& g1, Dual& g2) {
  const double m2 = mu.v * mu.v;
  double s1 = 0.0, d1 = 0.0, s2 = 0.0, d2 = 0.0;
  for (int j = (kRGTerms - 3) / 2; j >= 0; --j) {
    d1 = d1 * m2 + s1;
    s1 = s1 * m2 + kRG[2 * j + 2];
    d2 = d2 * m2 + s2;
    s2 = s2 * m2 + kRG[2 * j + 1];
  }
  // d/dmu f(mu^2) = f'(m2) * 2 mu
  g1 = Dual(-s1, -d1 * 2.0 * mu.v * mu.d);
  g2 = Dual(s2, d2 * 2.0 * mu.v * mu.d);
}

// Digamma for x > 0 (recurrence to x >= 10, then the asymptotic series).
inline double host_digamma(double x) {
  double r = 0.0;
  while (x < 10.0) {
    r -= 1.0 / x;
    x += 1.0;
  }
  const double f = 1.0 / (x * x);
  const double t = f * (-1.0 / 12 + f * (1.0 / 120 + f * (-1.0 / 252 + f * (1.0 / 240 + f * (-1.0 / 132 + f * (691.0 / 32760 + f * (-1.0 / 12)))))));
  return r + ::log(x) - 0.5 / x + t;
}

// Fill the per-order constants for order nu >= 0.  When with_derivative is
// set, the dual parts hold d/dnu (used only for KnuOrder<Dual>).
template <class T>
inline T host_make(Dual v);
template <>
inline double host_make<double>(Dual v) { return v.v; }
template <>
inline Dual host_make<Dual>(Dual v) { return v; }

template <class T>
inline void knu_prepare(double nu_in, KnuOrder<T>& o) {
  const double nuv = ::fabs(nu_in);
  const int n = static_cast<int>(::floor(nuv + 0.5));
  const Dual nu(nuv, 1.0);
  const Dual mu = nu - double(n);
  Dual g1, g2;
  host_temme_gammas(mu, g1, g2);
  const Dual rgp = host_rgamma1p(mu);
  const Dual rgm = host_rgamma1p(-mu);
  Dual fact;
  const double pm = M_PI * mu.v;
  if (::fabs(pm) < 1.0e-12) {
    fact = Dual(1.0, 0.0);
  } else {
    // d/dmu [pi mu / sin(pi mu)] = pi (sin - pm cos)/sin^2
    const double s = ::sin(pm), c = ::cos(pm);
    fact = Dual(pm / s, M_PI * (s - pm * c) / (s * s) * mu.d);
  }
  o.nu = host_make<T>(nu);
  o.mu = host_make<T>(mu);
  o.g1 = host_make<T>(g1);
  o.g2 = host_make<T>(g2);
  o.gp = host_make<T>(Dual(1.0) / rgp);
  o.gm = host_make<T>(Dual(1.0) / rgm);
  o.fact = host_make<T>(fact);
  o.a1 = host_make<T>(Dual(0.25) - mu * mu);
  for (int i = 1; i <= kTemmeTerms; ++i) {
    o.rden[i - 1] = host_make<T>(Dual(1.0) / (Dual(double(i) * i) - mu * mu));
    o.rmm[i - 1] = host_make<T>(Dual(1.0) / (Dual(double(i)) - mu));
    o.rmp[i - 1] = host_make<T>(Dual(1.0) / (Dual(double(i)) + mu));
  }
  o.n_rec = n;
  o.closed_form = 0;
  if (nuv == 0.5) o.closed_form = 1;
  if (nuv == 1.5) o.closed_form = 3;
  if (nuv == 2.5) o.closed_form = 5;
}

// 1/(2^(nu-1) Gamma(nu)) and its nu-derivative -con (ln 2 + psi(nu)).
inline Dual host_matern_con(double nu) {
  const double con = 1.0 / (::pow(2.0, nu - 1.0) * ::tgamma(nu));
  return Dual(con, -con * (M_LN2 + host_digamma(nu)));
}

// x^nu K_nu(x) / (2^(nu-1) Gamma(nu)) for x > 0 (value only); closed forms
// for nu = 1/2, 3/2, 5/2.
EXA_HD double matern_corr(const KnuOrder<double>& o, double con, double x) {
  if (o.closed_form == 1) return ::exp(-x);
  if (o.closed_form == 3) return (1.0 + x) * ::exp(-x);
  if (o.closed_form == 5) return (1.0 + x * (1.0 + x * (1.0 / 3.0))) * ::exp(-x);
  double k, k1;
  knu_pair(o, x, k, k1);
  return con * ::exp(o.nu * ::log(x)) * k;
}

}  // namespace exa
