#pragma once

namespace zolo {
double dlamch_p();
void mellipke(double alpha, double tol, double* k, double* e);
void mellipj(double u, double alpha, double tol, double* sn, double* cn, double* dn);
int choosem(double con);
void compute_coeff(double con, double tol, int rk, double* coeff);
void zolo_fone_con(double con, const double* coeff, int rk, double* fone, double* con_next);
double hfunc(double x);
}  // namespace zolo
