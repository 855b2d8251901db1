// Scalar math of the Zolotarev and QDWH iterations (FP64, host).
//
// mellipke / mellipj / choosem / compute_coeff restate Auxil_Zolo.f90:36-200,
// zolo_fone_con restates pzolopd1.f90:459-472 and hfunc restates hfunc.f90:27-39.
// They keep the reference's formulas and loop structure: the coefficients
// must match the reference's to rounding, because they select the iteration
// schedule (QR or Cholesky) and the partial-fraction weights.
#include "coeff.hpp"

#include <cfloat>
#include <cmath>
#include <vector>

namespace zolo {

double dlamch_p() { return DBL_EPSILON; }  // LAPACK DLAMCH('Precision') = 2^-52

void mellipke(double alpha, double tol, double* k, double* e) {
  const double m = std::sin(alpha) * std::sin(alpha);
  double a0 = 1.0, b0 = std::cos(alpha), s0 = m, mm = 1.0, a1 = a0;
  int i1 = 0;
  while (mm > tol) {
    a1 = (a0 + b0) / 2.0;
    const double b1 = std::sqrt(a0 * b0);
    const double c1 = (a0 - b0) / 2.0;
    ++i1;
    const double w1 = std::ldexp(1.0, i1) * c1 * c1;
    mm = w1;
    s0 += w1;
    a0 = a1;
    b0 = b1;
  }
  *k = M_PI / (2.0 * a1);
  *e = *k * (1.0 - s0 / 2.0);
}

void mellipj(double u, double alpha, double tol, double* sn, double* cn, double* dn) {
  const double m = std::sin(alpha) * std::sin(alpha);
  const int mchunk = 1000;
  std::vector<double> a(mchunk + 1), b(mchunk + 1), c(mchunk + 1), phin(mchunk + 2);
  a[1] = 1.0;
  b[1] = std::cos(alpha);
  c[1] = std::sin(alpha);
  int i = 1, nin = 0;
  while (std::fabs(c[i]) > tol && i < mchunk) {
    ++i;
    a[i] = (a[i - 1] + b[i - 1]) * 0.5;
    b[i] = std::sqrt(a[i - 1] * b[i - 1]);
    c[i] = (a[i - 1] - b[i - 1]) * 0.5;
    if (std::fabs(c[i]) <= tol && std::fabs(c[i - 1]) > tol) nin = i - 1;
  }
  phin[1] = u;
  phin[i] = std::ldexp(1.0, nin) * a[i] * u;
  while (i > 1) {
    --i;
    phin[i] = phin[i + 1];
    if (nin >= i) phin[i] = 0.5 * (std::asin(c[i + 1] * std::sin(phin[i + 1]) / a[i + 1]) + phin[i + 1]);
  }
  *sn = std::sin(phin[1]);
  *cn = std::cos(phin[1]);
  *dn = std::sqrt(1.0 - m * (*sn) * (*sn));
}

int choosem(double con) {
  if (con < 1.001) return 2;
  if (con <= 1.01) return 3;
  if (con <= 1.1) return 4;
  if (con <= 1.2) return 5;
  if (con <= 1.5) return 6;
  if (con <= 2) return 8;
  if (con < 6.5) return 2;
  if (con < 180) return 3;
  if (con < 1.5e4) return 4;
  if (con < 2e6) return 5;
  if (con < 1e9) return 6;
  if (con < 3e12) return 7;
  return 8;
}

void compute_coeff(double con, double tol, int rk, double* coeff) {
  const double kappa = 1.0 / con;
  const double alpha = std::acos(kappa);
  double kk, tmp;
  mellipke(alpha, tol, &kk, &tmp);
  const int rk2 = 2 * rk;
  for (int i = 1; i <= rk2; ++i) {
    const double t = i * kk / (rk2 + 1.0);
    double sn, cn, dn;
    mellipj(t, alpha, tol, &sn, &cn, &dn);
    coeff[i - 1] = sn * sn / (cn * cn);
  }
  for (int i = 1; i <= rk; ++i) {
    double v = 1.0;
    for (int j = 1; j <= rk; ++j) {
      v *= coeff[2 * i - 2] - coeff[2 * j - 1];
      if (i != j) v /= coeff[2 * i - 2] - coeff[2 * j - 2];
    }
    coeff[rk2 + i - 1] = v;
  }
}

void zolo_fone_con(double con, const double* coeff, int rk, double* fone, double* con_next) {
  double f = 1.0;
  for (int i = 1; i <= rk; ++i) f *= (1.0 + coeff[2 * i - 1]) / (1.0 + coeff[2 * i - 2]);
  double fc = con;
  for (int i = 1; i <= rk; ++i) fc *= (con * con + coeff[2 * i - 1]) / (con * con + coeff[2 * i - 2]);
  *fone = f;
  *con_next = std::fmax(fc / f, 1.0);
}

double hfunc(double x) {
  const double tol = 30 * dlamch_p();
  const double x2 = x * x;
  double ww = 0.0;
  if (!(std::fabs(x - 1.0) <= tol)) {
    const double tmp = 4.0 * (1.0 - x2) / (x2 * x2);
    ww = std::exp(std::log(tmp) / 3.0);
  }
  const double sum1 = std::sqrt(1.0 + ww);
  const double h = std::sqrt(8.0 - 4.0 * ww + 8.0 * (2.0 - x2) / (x2 * sum1));
  return sum1 + h / 2.0;
}

}  // namespace zolo
