// MatrixMarket input assembled on the GPU: the de-duplicated triplets are
// uploaded, scattered into the zeroed matrix and, for symmetric files,
// symmetrised with the diagonal shift (the same values, bit for bit, as the
// host assembly of mm_reader.cpp; prepare_matrix.f90:638-759).
#include <vector>

#include "kernels.hpp"
#include "mm_reader.hpp"

namespace zolo {

int mm_dense_device(const char* path, double sigma, double* A, int64_t lda, int* mtype, hipStream_t s) {
  MmFile m;
  ZOLO_CHECK(mm_parse(path, true, m));
  ZOLO_CHECK(mm_check_dense(m, A, lda));
  mm_dedupe_last(m);
  const int64_t n = m.rows, z = (int64_t)m.v.size();
  ZOLO_CHECK(k_fill<double>(A, lda, 0, n, n, 0.0, 1, s));
  if (z > 0) {
    // triplets: 0-based row and column (int64) and the value
    std::vector<int64_t> ij((size_t)2 * z);
    for (int64_t k = 0; k < z; ++k) {
      ij[(size_t)k] = m.i[(size_t)k] - 1;
      ij[(size_t)(z + k)] = m.j[(size_t)k] - 1;
    }
    char* d = nullptr;
    const size_t bytes = sizeof(int64_t) * 2 * (size_t)z + sizeof(double) * (size_t)z;
    ZOLO_HIP(hipMalloc(&d, bytes));
    struct Free {
      char* p;
      ~Free() { (void)hipFree(p); }
    } fr{d};
    int64_t* dij = reinterpret_cast<int64_t*>(d);
    double* dv = reinterpret_cast<double*>(d + sizeof(int64_t) * 2 * (size_t)z);
    ZOLO_HIP(hipMemcpyAsync(dij, ij.data(), sizeof(int64_t) * 2 * (size_t)z, hipMemcpyHostToDevice, s));
    ZOLO_HIP(hipMemcpyAsync(dv, m.v.data(), sizeof(double) * (size_t)z, hipMemcpyHostToDevice, s));
    ZOLO_CHECK(k_scatter(A, lda, dij, dij + z, dv, z, s));
    // the host arrays and the device triplets must outlive the copies and the scatter
    ZOLO_HIP(hipStreamSynchronize(s));
  }
  if (m.mtype == 0) ZOLO_CHECK(k_mm_symmetrize(A, lda, n, sigma, s));
  if (mtype) *mtype = m.mtype;
  return kOk;
}

}  // namespace zolo
