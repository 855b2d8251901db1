// MatrixMarket input: the coordinate files the reference's drivers read
// (mmread, prepare_matrix.f90:1076-1399) assembled into a dense column-major
// matrix as prepare_real_GEM (prepare_matrix.f90:154-321) and
// prepare_real_symm_MM (:638-759) do:
//
//  * every entry (i, j, v) is stored as A(i, j) = v in file order (a repeated
//    position keeps the last value);
//  * 'symmetric' files: A <- A + A^T, then the diagonal A(i, i) <- A(i, i)/2 -
//    sigma, so each stored triangle entry appears on both sides and the
//    diagonal is shifted by -sigma;
//  * 'general', 'skew-symmetric' and 'hermitian' files: A as stored, no shift
//    (the reference's mtype 1 and 2).
//
// Fields 'real', 'integer' and 'pattern' (value 1) are read; 'complex' and the
// dense 'array' form are refused (the reference's assembly uses only the
// real values and the row/column indices of coordinate files).  Parsing and
// the host assembly are plain host code (no HIP calls); the device assembly
// is in mm_device.cpp.
#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "mm_reader.hpp"

namespace zolo {
namespace {

std::string lower(std::string s) {
  for (auto& ch : s) ch = (char)std::tolower((unsigned char)ch);
  return s;
}

// One line without its newline; false at end of file.
bool get_line(FILE* f, std::string& out) {
  out.clear();
  char buf[4096];
  bool any = false;
  while (std::fgets(buf, sizeof buf, f)) {
    any = true;
    out += buf;
    if (!out.empty() && out.back() == '\n') {
      out.pop_back();
      if (!out.empty() && out.back() == '\r') out.pop_back();
      return true;
    }
  }
  return any;
}

bool blank(const std::string& s) {
  for (char ch : s)
    if (!std::isspace((unsigned char)ch)) return false;
  return true;
}

// Fortran list-directed input also takes D exponents (1.5D+02).
double parse_real(const char* p, char** end) {
  char tmp[128];
  size_t k = 0;
  while (p[k] && !std::isspace((unsigned char)p[k]) && k + 1 < sizeof tmp) {
    tmp[k] = (p[k] == 'd' || p[k] == 'D') ? 'e' : p[k];
    ++k;
  }
  tmp[k] = 0;
  char* e = nullptr;
  const double x = std::strtod(tmp, &e);
  *end = (e == tmp) ? const_cast<char*>(p) : const_cast<char*>(p) + (e - tmp);
  return x;
}

int mm_error(int code, const std::string& msg) {
  set_error("MatrixMarket: " + msg);
  return code;
}

}  // namespace

// Header and size line; with entries = true also every entry.
int mm_parse(const char* path, bool entries, MmFile& m) {
  if (!path) return mm_error(-1, "no path");
  FILE* f = std::fopen(path, "r");
  if (!f) return mm_error(-1, std::string("cannot open ") + path);
  struct Closer {
    FILE* f;
    ~Closer() { std::fclose(f); }
  } closer{f};
  std::string line;
  if (!get_line(f, line)) return mm_error(-2, "empty file");
  char w[5][64] = {{0}};
  const int nw = std::sscanf(line.c_str(), "%63s %63s %63s %63s %63s", w[0], w[1], w[2], w[3], w[4]);
  if (nw < 5 || lower(w[0]) != "%%matrixmarket") return mm_error(-2, "missing %%MatrixMarket banner");
  const std::string object = lower(w[1]), rep = lower(w[2]), field = lower(w[3]), symm = lower(w[4]);
  if (object != "matrix") return mm_error(-2, "object '" + object + "' is not 'matrix'");
  if (rep != "coordinate") return mm_error(-2, "format '" + rep + "' is not 'coordinate'");
  if (field != "real" && field != "integer" && field != "pattern")
    return mm_error(-2, "field '" + field + "' is not real, integer or pattern");
  if (symm == "symmetric") m.mtype = 0;
  else if (symm == "general") m.mtype = 1;
  else if (symm == "skew-symmetric" || symm == "hermitian") m.mtype = 2;
  else return mm_error(-2, "symmetry '" + symm + "' not recognised");
  m.pattern = field == "pattern";
  // comments and blank lines, then the size line
  for (;;) {
    if (!get_line(f, line)) return mm_error(-3, "no size line");
    if (!line.empty() && line[0] == '%') continue;
    if (blank(line)) continue;
    break;
  }
  long long r = 0, c = 0, z = 0;
  if (std::sscanf(line.c_str(), "%lld %lld %lld", &r, &c, &z) != 3 || r < 0 || c < 0 || z < 0)
    return mm_error(-3, "bad size line '" + line + "'");
  m.rows = r; m.cols = c; m.nnz = z;
  if (!entries) return kOk;
  // An entry line takes at least 4 bytes ("i j" and a newline), so a size line
  // that claims more entries than the rest of the file can hold is refused
  // before anything is allocated; storage then grows with the entries read.
  const long here = std::ftell(f);
  if (here >= 0 && std::fseek(f, 0, SEEK_END) == 0) {
    const long end = std::ftell(f);
    if (end < here || std::fseek(f, here, SEEK_SET) != 0) return mm_error(-3, "cannot size the file");
    if (z > (long long)(end - here) / 4 + 1)
      return mm_error(-3, "size line claims " + std::to_string(z) + " entries, more than the file holds");
  }
  m.i.reserve((size_t)std::min<long long>(z, 1 << 20));
  m.j.reserve(m.i.capacity());
  m.v.reserve(m.i.capacity());
  int64_t k = 0;
  while (k < z && get_line(f, line)) {
    if (blank(line) || line[0] == '%') continue;
    const char* p = line.c_str();
    char* e = nullptr;
    const long long ii = std::strtoll(p, &e, 10);
    if (e == p) return mm_error(-4, "bad entry line '" + line + "'");
    p = e;
    const long long jj = std::strtoll(p, &e, 10);
    if (e == p) return mm_error(-4, "bad entry line '" + line + "'");
    p = e;
    double val = 1.0;
    if (!m.pattern) {
      while (*p && std::isspace((unsigned char)*p)) ++p;
      val = parse_real(p, &e);
      if (e == p) return mm_error(-4, "bad entry value in '" + line + "'");
    }
    if (ii < 1 || ii > r || jj < 1 || jj > c)
      return mm_error(-5, "entry (" + std::to_string(ii) + ", " + std::to_string(jj) + ") outside " +
                              std::to_string(r) + " x " + std::to_string(c));
    m.i.push_back(ii);
    m.j.push_back(jj);
    m.v.push_back(val);
    ++k;
  }
  if (k < z) return mm_error(-4, "file ends after " + std::to_string(k) + " of " + std::to_string(z) + " entries");
  return kOk;
}

// Keep the last entry of every repeated position (file order decides, as the
// reference's sequential stores); result sorted by column, then row.
void mm_dedupe_last(MmFile& m) {
  const size_t z = m.v.size();
  std::vector<size_t> idx(z);
  for (size_t k = 0; k < z; ++k) idx[k] = k;
  std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) {
    return m.j[a] != m.j[b] ? m.j[a] < m.j[b] : m.i[a] < m.i[b];
  });
  std::vector<int64_t> i2, j2;
  std::vector<double> v2;
  i2.reserve(z); j2.reserve(z); v2.reserve(z);
  for (size_t k = 0; k < z; ++k) {
    const size_t a = idx[k];
    if (k + 1 < z && m.i[idx[k + 1]] == m.i[a] && m.j[idx[k + 1]] == m.j[a]) continue;  // a later one follows
    i2.push_back(m.i[a]);
    j2.push_back(m.j[a]);
    v2.push_back(m.v[a]);
  }
  m.i.swap(i2); m.j.swap(j2); m.v.swap(v2);
}

int mm_check_dense(const MmFile& m, double* A, int64_t lda) {
  if (m.rows != m.cols) return mm_error(-6, "matrix is not square");
  if (!A) return mm_error(-7, "no output array");
  if (lda < std::max<int64_t>(m.rows, 1)) return mm_error(-8, "lda smaller than the order");
  return kOk;
}

int mm_info(const char* path, int64_t* rows, int64_t* cols, int64_t* nnz, int* mtype, int64_t* bandwidth) {
  MmFile m;
  ZOLO_CHECK(mm_parse(path, bandwidth != nullptr, m));
  if (rows) *rows = m.rows;
  if (cols) *cols = m.cols;
  if (nnz) *nnz = m.nnz;
  if (mtype) *mtype = m.mtype;
  if (bandwidth) {  // max |i - j| + 1 over the entries (prepare_matrix.f90:223-224)
    int64_t bw = 0;
    for (size_t k = 0; k < m.v.size(); ++k) bw = std::max<int64_t>(bw, std::llabs(m.i[k] - m.j[k]));
    *bandwidth = m.v.empty() ? 0 : bw + 1;
  }
  return kOk;
}

int mm_dense_host(const char* path, double sigma, double* A, int64_t lda, int* mtype) {
  MmFile m;
  ZOLO_CHECK(mm_parse(path, true, m));
  ZOLO_CHECK(mm_check_dense(m, A, lda));
  const int64_t n = m.rows;
  for (int64_t c = 0; c < n; ++c) std::fill(A + c * lda, A + c * lda + n, 0.0);
  for (size_t k = 0; k < m.v.size(); ++k) A[(m.i[k] - 1) + (m.j[k] - 1) * lda] = m.v[k];
  if (m.mtype == 0) {
    // A <- A + A^T (PDTRAN with beta = 1), then A(i, i) <- A(i, i)/2 - sigma
    for (int64_t c = 0; c < n; ++c)
      for (int64_t r = c + 1; r < n; ++r) {
        const double s = A[r + c * lda] + A[c + r * lda];
        A[r + c * lda] = s;
        A[c + r * lda] = s;
      }
    for (int64_t d = 0; d < n; ++d) A[d + d * lda] = (A[d + d * lda] + A[d + d * lda]) / 2.0 - sigma;
  }
  if (mtype) *mtype = m.mtype;
  return kOk;
}

}  // namespace zolo
