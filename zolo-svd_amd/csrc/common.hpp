// Shared definitions for the Zolo-PD HIP library (gfx950 / CDNA4).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>

namespace zolo {

// Status codes returned through the C ABI (mirrors include/zolo.h).
enum Status : int {
  kOk = 0,
  kErrArg = -1,
  kErrHip = 1001,
  kErrNotSpd = 1002,
  kErrNoConvergence = 1003,
  kErrAlloc = 1004,
  kErrInternal = 1005,
};

// Thread-local last error message, readable through zolo_last_error().
void set_error(const std::string& msg);
const char* last_error();

#define ZOLO_HIP(expr)                                                        \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess) {                                                   \
      ::zolo::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));   \
      return ::zolo::kErrHip;                                                 \
    }                                                                         \
  } while (0)

#define ZOLO_CHECK(...)                                                       \
  do {                                                                        \
    int s_ = (__VA_ARGS__);                                                   \
    if (s_ != 0) return s_;                                                   \
  } while (0)

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// ---------------------------------------------------------------------------
// GEMM descriptor.  Column-major operands, C = alpha * op(A) op(B) + beta * C,
// op(A) is M x K, op(B) is K x N.  Batched over blockIdx.z with element strides.
// K-range trimming: the caller may declare that op(A) / op(B) are exactly zero
// outside a band, so that tiles skip the zero part of the K loop:
//   aLo >= 0 : op(A)[i][k] == 0 for k < i - aLo     (upper-type)
//   aHi >= 0 : op(A)[i][k] == 0 for k > i + aHi     (lower-type)
//   bLo >= 0 : op(B)[k][j] == 0 for k < j - bLo     (lower-type in (k,j))
//   bHi >= 0 : op(B)[k][j] == 0 for k > j + bHi     (upper-type in (k,j))
// The zeros must be present in memory; trimming is an optimisation only.
// lowerC: only tiles intersecting the lower triangle (i >= j) are computed.
// ---------------------------------------------------------------------------
struct GemmDesc {
  int transA = 0, transB = 0;
  int64_t M = 0, N = 0, K = 0;
  const void* A = nullptr; int64_t lda = 0, strideA = 0;
  const void* B = nullptr; int64_t ldb = 0, strideB = 0;
  void* C = nullptr; int64_t ldc = 0, strideC = 0;
  double alpha = 1.0, beta = 0.0;
  const double* alphaVec = nullptr;   // optional per-batch alpha (device), multiplies alpha
  int batch = 1;
  int aLo = -1, aHi = -1, bLo = -1, bHi = -1;
  int lowerC = 0;
  // optional device tables of the batch index of each operand (map[z] instead
  // of z): a batch over (matrix, shift) items reads the one X of its matrix
  const int* mapA = nullptr;
  const int* mapB = nullptr;
  const int* mapC = nullptr;
  // tile hint: 0 auto, 1 128x128, 2 64x128, 3 128x64, 4 64x64 (F64 / F32)
  int tile = 0;
  // F64X: the caller guarantees that op(A) holds the same values as in the
  // previous F64X product (same pointer and layout, not written since), so
  // its residues may be reused when that product left them in the workspace
  int sameA = 0;
  // Two row ranges (F64 / F32 engine): rows k >= bGapAt of B (transB = 0)
  // are read from row k + bGap, and rows i >= cGapAt of C are read and
  // written at row i + cGap, so a product over [rows r0, r1) of one block and
  // [n, r2) of another is one launch.  bGapAt: a multiple of 16 with no lower
  // K trimming (aLo, bLo < 0); cGapAt: a multiple of 128; -1: no gap.
  int64_t bGapAt = -1, bGap = 0;
  int64_t cGapAt = -1, cGap = 0;
  // split-K (set by the launcher): blockIdx.y selects a K chunk whose raw
  // partial product goes to C + blockIdx.y * splitStride
  int splitk = 1;
  int64_t kchunk = 0;
  int64_t splitStride = 0;
};

// F64X: FP64 operands and result, the products formed exactly from int8
// residues modulo 16 primes-like moduli (ozaki.hip); FP64-level accuracy.
enum class Prec : int { F64 = 0, F32 = 1, BF16X3 = 2, F64X = 3 };

int gemm(Prec p, const GemmDesc& d, hipStream_t s);
// The int8-modular FP64 product (ozaki.hip); descriptors it does not cover
// (operand maps, K > 131072) take the FP64 MFMA engine in gemm().
bool gemm_ozaki_supported(const GemmDesc& d);
int gemm_ozaki(const GemmDesc& d, hipStream_t s, bool capture);
// Graph-capture mode: no split-K (its workspace may need an allocation) and
// no launch ledger events while on.
void gemm_set_capture(bool on);

// GEMM launch ledger (see gemm.hip).
void ledger_enable(bool on);
int ledger_query(int prec, int ta, int tb, int tm, int tn, int64_t* count, double* ms, double* flops);

}  // namespace zolo
