// Host-side checks for the AddressSanitizer / UndefinedBehaviorSanitizer
// build (make -C zolo-svd_amd sanitize): the MatrixMarket reader on valid and
// malformed files and the scalar Zolotarev math (Auxil_Zolo.f90:36-200,
// hfunc.f90:2-42) over a range of arguments.  Plain host code, compiled with
// g++; it needs no GPU.  Usage: host_checks <scratch directory>.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "coeff.hpp"
#include "mm_reader.hpp"

namespace zolo {
int mm_info(const char* path, int64_t* rows, int64_t* cols, int64_t* nnz, int* mtype, int64_t* bandwidth);
int mm_dense_host(const char* path, double sigma, double* A, int64_t lda, int* mtype);
}  // namespace zolo

namespace {

int failures = 0;

void expect(bool ok, const std::string& what) {
  if (!ok) {
    std::printf("FAIL: %s\n", what.c_str());
    ++failures;
  }
}

std::string write_file(const std::string& dir, const std::string& name, const std::string& text) {
  const std::string path = dir + "/" + name;
  FILE* f = std::fopen(path.c_str(), "w");
  if (!f) return path;
  std::fwrite(text.data(), 1, text.size(), f);
  std::fclose(f);
  return path;
}

// status of mm_dense_host on a file (and the assembled matrix for n <= 8)
int read_dense(const std::string& path, std::vector<double>& A, int* mtype) {
  int64_t r = 0, c = 0, z = 0;
  int t = -1;
  const int st = zolo::mm_info(path.c_str(), &r, &c, &z, &t, nullptr);
  if (st != 0) return st;
  const int64_t n = std::max<int64_t>(r, 1);
  A.assign((size_t)n * n, -7.0);
  return zolo::mm_dense_host(path.c_str(), 0.25, A.data(), n, mtype);
}

void check_matrix_market(const std::string& dir) {
  std::vector<double> A;
  int t = -1;
  // symmetric with a D exponent: A + A^T, diagonal halved and shifted by -0.25
  std::string p = write_file(dir, "sym.mtx",
                             "%%MatrixMarket matrix coordinate real symmetric\n% comment\n\n3 3 4\n1 1 2.0\n2 1 -1.5D+00\n"
                             "3 3 4\n3 2 0.5\n");
  expect(read_dense(p, A, &t) == 0 && t == 0, "symmetric file reads");
  expect(A.size() == 9 && A[0] == 1.75 && A[1] == -1.5 && A[3] == -1.5 && A[8] == 3.75 && A[5] == 0.5 && A[7] == 0.5,
         "symmetric assembly");
  // general, pattern, integer, repeated position (last value wins)
  p = write_file(dir, "gen.mtx", "%%MatrixMarket matrix coordinate integer general\n2 2 3\n1 2 5\n2 1 -3\n1 2 7\n");
  expect(read_dense(p, A, &t) == 0 && t == 1 && A[2] == 7.0 && A[1] == -3.0 && A[0] == 0.0, "general integer file");
  p = write_file(dir, "pat.mtx", "%%MatrixMarket matrix coordinate pattern general\n2 2 2\n1 1\n2 2\n");
  expect(read_dense(p, A, &t) == 0 && A[0] == 1.0 && A[3] == 1.0 && A[1] == 0.0, "pattern file");
  int64_t r = 0, c = 0, z = 0, bw = 0;
  expect(zolo::mm_info(p.c_str(), &r, &c, &z, &t, &bw) == 0 && bw == 1, "bandwidth of a diagonal file");
  // malformed files: each returns its code, none reads or writes out of bounds
  struct Bad {
    const char* name;
    const char* text;
    int code;
  } bad[] = {
      {"empty.mtx", "", -2},
      {"banner.mtx", "%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n", -2},
      {"field.mtx", "%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1 0\n", -2},
      {"nosize.mtx", "%%MatrixMarket matrix coordinate real general\n% only a comment\n", -3},
      {"size.mtx", "%%MatrixMarket matrix coordinate real general\n2 -2 1\n1 1 1\n", -3},
      {"huge.mtx", "%%MatrixMarket matrix coordinate real general\n4 4 9000000000000000000\n1 1 1\n", -3},
      {"short.mtx", "%%MatrixMarket matrix coordinate real general\n3 3 3\n1 1 1\n", -3},
      {"early.mtx", "%%MatrixMarket matrix coordinate real general\n3 3 2\n1 1 1\n% x\n% y\n", -4},
      {"entry.mtx", "%%MatrixMarket matrix coordinate real general\n3 3 1\nx 1 1\n", -4},
      {"value.mtx", "%%MatrixMarket matrix coordinate real general\n3 3 1\n1 1 abc\n", -4},
      {"range.mtx", "%%MatrixMarket matrix coordinate real general\n3 3 1\n4 1 1.0\n", -5},
      {"zero.mtx", "%%MatrixMarket matrix coordinate real general\n3 3 1\n0 1 1.0\n", -5},
      {"rect.mtx", "%%MatrixMarket matrix coordinate real general\n3 2 1\n1 1 1.0\n", -6},
  };
  for (const Bad& b : bad) {
    p = write_file(dir, b.name, b.text);
    std::vector<double> B;
    const int st = read_dense(p, B, &t);
    expect(st == b.code, std::string(b.name) + " gives " + std::to_string(st) + ", expected " +
                             std::to_string(b.code));
  }
  expect(zolo::mm_info((dir + "/missing.mtx").c_str(), &r, &c, &z, &t, nullptr) == -1, "missing file");
  expect(zolo::mm_info(nullptr, &r, &c, &z, &t, nullptr) == -1, "no path");
  // a value field longer than the reader's token buffer and an entry line
  // longer than its line buffer
  std::string longv = "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1 1.";
  longv += std::string(300, '5');
  longv += "\n";
  p = write_file(dir, "longvalue.mtx", longv);
  expect(read_dense(p, A, &t) == 0 && std::fabs(A[0] - 1.5555555) < 1e-6, "long value token");
  std::string longl = "%%MatrixMarket matrix coordinate real general\n2 2 1\n2 2 3.0 ";
  longl += std::string(10000, ' ');
  longl += "\n";
  p = write_file(dir, "longline.mtx", longl);
  expect(read_dense(p, A, &t) == 0 && A[3] == 3.0, "line longer than the read buffer");
  // lda smaller than the order, no output array
  p = write_file(dir, "ok.mtx", "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1 1.0\n");
  double one[4];
  expect(zolo::mm_dense_host(p.c_str(), 0.0, one, 1, &t) == -8, "lda < n");
  expect(zolo::mm_dense_host(p.c_str(), 0.0, nullptr, 2, &t) == -7, "no output array");
}

void check_scalars() {
  const double eps = 2.220446049250313e-16;
  const double cons[] = {1.0, 1.0000001, 1.5, 3.0, 6.5, 180.0, 1.5e4, 2e6, 1e9, 3e12, 1e15, 1e18};
  for (double con : cons) {
    const int r = zolo::choosem(con);
    expect(r >= 1 && r <= 8, "choosem in range");
    for (int rk = 1; rk <= 8; ++rk) {
      double coeff[24];
      zolo::compute_coeff(con, eps, rk, coeff);
      bool finite = true;
      for (int i = 0; i < 3 * rk; ++i) finite = finite && std::isfinite(coeff[i]);
      expect(finite || con == 1.0, "coefficients finite at con " + std::to_string(con));
      double fone = 0.0, cn = 0.0;
      zolo::zolo_fone_con(con, coeff, rk, &fone, &cn);
      expect(con == 1.0 || (cn >= 1.0 && cn <= con), "CON recurrence contracts");
    }
  }
  // QDWH's lower bounds lie in [u, 1] (the formula's l^4 underflows below ~1e-75)
  for (double x : {1e-20, 1e-12, 0.3, 1.0, 1.0 - 1e-15}) expect(std::isfinite(zolo::hfunc(x)), "hfunc finite");
  double k = 0, e = 0, sn = 0, cn = 0, dn = 0;
  for (double a : {0.0, 0.1, 1.0, 1.5707963267948966}) {
    zolo::mellipke(a, eps, &k, &e);
    expect(std::isfinite(k) || a > 1.57, "mellipke finite");
    zolo::mellipj(0.3, a, eps, &sn, &cn, &dn);
    expect(std::isfinite(sn) && std::isfinite(cn) && std::isfinite(dn), "mellipj finite");
  }
}

}  // namespace

int main(int argc, char** argv) {
  const std::string dir = argc > 1 ? argv[1] : ".";
  check_matrix_market(dir);
  check_scalars();
  if (failures == 0) std::printf("host checks passed\n");
  return failures == 0 ? 0 : 1;
}
