// MatrixMarket coordinate files (mmread, prepare_matrix.f90:1076-1399): the
// parsed entries and the host-side steps shared by the host and the device
// assembly.  Host code only.
#pragma once
#include <cstdint>
#include <vector>

#include "common.hpp"

namespace zolo {

struct MmFile {
  int64_t rows = 0, cols = 0, nnz = 0;
  int mtype = 2;   // 0 symmetric, 1 general, 2 other (prepare_matrix.f90:230-241)
  bool pattern = false;
  std::vector<int64_t> i, j;  // 1-based as in the file
  std::vector<double> v;
};

// Header and size line; with entries = true also every entry.  Status 0 or
// the codes of zolo_mm_info (-1 .. -5).
int mm_parse(const char* path, bool entries, MmFile& m);
// Keep the last entry of every repeated position; sorted by column, then row.
void mm_dedupe_last(MmFile& m);
// The square / output-array checks of the dense assembly (-6 .. -8).
int mm_check_dense(const MmFile& m, double* A, int64_t lda);

}  // namespace zolo
