// LDS-tiled MFMA GEMM engine for gfx950 (CDNA4), column-major operands.
//
// One engine serves every dense contraction of the polar-decomposition path:
// the Gram matrix X^T X (lower triangle only), the Cholesky trailing updates,
// the triangular multiplies by explicit inverse factors, the compact-WY
// updates of the structured Householder QR and the Q1 Q2^T product.
//
//  * F64    : v_mfma_f64_16x16x4f64, 128x128x16 tile, 4 waves of 64x64.
//  * F32    : v_mfma_f32_16x16x4f32, same tiling.
//  * BF16X3 : fp32 operands split at LDS staging into hi + lo bfloat16 halves,
//             three v_mfma_f32_16x16x32_bf16 products (hi*hi + hi*lo + lo*hi)
//             accumulated in fp32; 128x128x32 tile.
//
// Work distribution: tiles are numbered so that the 8 XCDs (round-robin
// workgroup dispatch) each own a contiguous block of tiles, and inside a block
// tiles walk down groups of 8 tile-rows so neighbouring workgroups share the
// same A and B panels in their XCD's L2.
#include <algorithm>
#include <algorithm>
#include <vector>

#include "common.hpp"

namespace zolo {

// ---------------------------------------------------------------------------
// Launch ledger: while enabled, every GEMM launch is bracketed by hipEvents on
// its own stream and its algorithmic flop count is recorded, so a caller can
// measure the engine's average launch duration over a real run.
// ---------------------------------------------------------------------------
namespace {
struct LedgerEntry {
  int prec, ta, tb, tm, tn;
  double flops;
  hipEvent_t e0, e1;
};
bool g_ledger_on = false;
std::vector<LedgerEntry> g_ledger;

// Number of (r, k) pairs, r < R, k < K, inside the band lo/hi (k >= r - lo,
// k <= r + hi; a negative offset means no bound on that side).
double band_pairs(int64_t R, int64_t K, int lo, int hi) {
  double n = 0.0;
  for (int64_t r = 0; r < R; ++r) {
    const int64_t kb = lo >= 0 ? std::max<int64_t>(0, r - lo) : 0;
    const int64_t ke = hi >= 0 ? std::min<int64_t>(K, r + hi + 1) : K;
    if (ke > kb) n += (double)(ke - kb);
  }
  return n;
}

double algorithmic_flops(const GemmDesc& d) {
  // 2 M N K scaled by the nonzero fraction of each banded operand (exact for
  // the triangular and trapezoidal operands of the path) and by the computed
  // fraction of C for lower-triangle-only outputs.
  double f = 2.0 * (double)d.M * (double)d.N * (double)d.K * d.batch;
  if (d.K > 0 && (d.aLo >= 0 || d.aHi >= 0)) f *= band_pairs(d.M, d.K, d.aLo, d.aHi) / ((double)d.M * d.K);
  if (d.K > 0 && (d.bLo >= 0 || d.bHi >= 0)) f *= band_pairs(d.N, d.K, d.bLo, d.bHi) / ((double)d.N * d.K);
  if (d.lowerC) f *= 0.5 * (double)(d.N + 1) / (double)d.N;
  return f;
}
}  // namespace

void ledger_enable(bool on) {
  for (auto& e : g_ledger) {
    (void)hipEventDestroy(e.e0);
    (void)hipEventDestroy(e.e1);
  }
  g_ledger.clear();
  g_ledger_on = on;
}

// Sum over recorded launches matching (prec, ta, tb, tm, tn); -1 matches any.
int ledger_query(int prec, int ta, int tb, int tm, int tn, int64_t* count, double* ms, double* flops) {
  *count = 0;
  *ms = 0.0;
  *flops = 0.0;
  for (auto& e : g_ledger) {
    if ((prec >= 0 && e.prec != prec) || (ta >= 0 && e.ta != ta) || (tb >= 0 && e.tb != tb) ||
        (tm >= 0 && e.tm != tm) || (tn >= 0 && e.tn != tn))
      continue;
    float t = 0.f;
    ZOLO_HIP(hipEventSynchronize(e.e1));
    ZOLO_HIP(hipEventElapsedTime(&t, e.e0, e.e1));
    *count += 1;
    *ms += t;
    *flops += e.flops;
  }
  return kOk;
}

namespace {

constexpr int kThreads = 256;
constexpr int BM = 128, BN = 128;   // bf16x3 tile

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

template <typename T> struct Acc;
template <> struct Acc<double> { typedef d4 type; };
template <> struct Acc<float> { typedef f4 type; };

__device__ inline d4 mfma16(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ inline f4 mfma16(float a, float b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Workgroup barrier that first waits for this wave's LDS operations only:
// global loads in flight stay in flight across it.  The memory clobber keeps
// the compiler from moving LDS accesses across the barrier.
__device__ inline void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

struct TileCoord { int64_t i0, j0; bool skip; };

// Map the linear workgroup id to a (row, col) tile, XCD-aware and grouped.
template <int BM, int BN>
__device__ inline TileCoord tile_of(int64_t M, int64_t N, int lowerC) {
  const int64_t tm = (M + BM - 1) / BM, tn = (N + BN - 1) / BN;
  const int64_t nt = tm * tn;
  int64_t pid = blockIdx.x;
  const int64_t full = (nt / 8) * 8;
  if (pid < full) {  // XCD x (= pid % 8) owns logical tiles [x*nt/8, (x+1)*nt/8)
    pid = (pid % 8) * (full / 8) + pid / 8;
  }
  const int64_t G = 8;  // tile-rows per group
  const int64_t group = pid / (G * tn);
  const int64_t first = group * G;
  const int64_t gsz = (tm - first) < G ? (tm - first) : G;
  const int64_t r = pid - group * G * tn;
  TileCoord t;
  t.i0 = (first + r % gsz) * BM;
  t.j0 = (r / gsz) * BN;
  t.skip = lowerC && (t.j0 > t.i0 + BM - 1);
  return t;
}

template <int BM, int BN>
__device__ inline void k_range(const GemmDesc& d, int64_t i0, int64_t j0, int64_t& kb, int64_t& ke) {
  kb = 0;
  ke = d.K;
  if (d.aLo >= 0) kb = max(kb, i0 - d.aLo);
  if (d.bLo >= 0) kb = max(kb, j0 - d.bLo);
  if (d.aHi >= 0) ke = min(ke, i0 + BM + d.aHi);
  if (d.bHi >= 0) ke = min(ke, j0 + BN + d.bHi);
}

// ---------------------------------------------------------------------------
// F64 / F32 kernel.  Tile TBM x TBN x 16, LDS double-buffered; the four waves form a WM x WN grid,
// each wave owning FM x FN blocks of 16x16 MFMA outputs.
// ---------------------------------------------------------------------------
template <typename T, int TA, int TB, int TBM, int TBN>
__global__ __launch_bounds__(kThreads) void gemm_kernel(GemmDesc d) {
  constexpr int BK = 16;
  constexpr int PAD = (sizeof(T) == 8) ? 4 : 8;
  constexpr int WM = (TBM == TBN) ? 2 : (TBM == 64 ? 1 : 4);
  constexpr int WN = 4 / WM;
  constexpr int FM = TBM / WM / 16, FN = TBN / WN / 16;
  constexpr int EA = TBM * BK / kThreads;   // A elements per thread
  constexpr int EB = TBN * BK / kThreads;   // B elements per thread
  __shared__ T As[2][BK][TBM + PAD];
  __shared__ T Bs[2][BK][TBN + PAD];

  const TileCoord tc = tile_of<TBM, TBN>(d.M, d.N, d.lowerC);
  if (tc.skip) return;
  const int64_t i0 = tc.i0, j0 = tc.j0;
  const int64_t z = blockIdx.z;
  const T* __restrict__ A = reinterpret_cast<const T*>(d.A) + (d.mapA ? (int64_t)d.mapA[z] : z) * d.strideA;
  const T* __restrict__ B = reinterpret_cast<const T*>(d.B) + (d.mapB ? (int64_t)d.mapB[z] : z) * d.strideB;
  T* __restrict__ C = reinterpret_cast<T*>(d.C) + (d.mapC ? (int64_t)d.mapC[z] : z) * d.strideC;

  int64_t kb, ke;
  k_range<TBM, TBN>(d, i0, j0, kb, ke);
  if (d.splitk > 1) {
    const int64_t y = blockIdx.y;
    kb = max(kb, y * d.kchunk);
    ke = min(ke, (y + 1) * d.kchunk);
    C += y * d.splitStride;
  }

  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int wm = wave / WN, wn = wave % WN;

  typedef typename Acc<T>::type AccT;
  AccT acc[FM][FN];
#pragma unroll
  for (int a = 0; a < FM; ++a)
#pragma unroll
    for (int b = 0; b < FN; ++b) acc[a][b] = AccT{0, 0, 0, 0};

  T ra[EA], rb[EB];
  // A: TA==0 -> EA consecutive i at fixed k; TA==1 -> EA consecutive k at fixed i.
  const int a_r = (TA == 0) ? t / (TBM / EA) : t / (BK / EA);
  const int a_c = (TA == 0) ? (t % (TBM / EA)) * EA : (t % (BK / EA)) * EA;
  // B: TB==0 -> EB consecutive k at fixed j; TB==1 -> EB consecutive j at fixed k.
  const int b_r = (TB == 0) ? t / (BK / EB) : t / (TBN / EB);
  const int b_c = (TB == 0) ? (t % (BK / EB)) * EB : (t % (TBN / EB)) * EB;

  // Interior tiles take unguarded loads (no per-element exec branches); edge
  // tiles and the K remainder take the guarded path.
  const bool interior = (i0 + TBM <= d.M) && (j0 + TBN <= d.N);
  auto load_fast = [&](int64_t k0) {
    if (TA == 0) {
      const T* src = A + (i0 + a_c) + (k0 + a_r) * d.lda;
#pragma unroll
      for (int e = 0; e < EA; ++e) ra[e] = src[e];
    } else {
      const T* src = A + (k0 + a_c) + (i0 + a_r) * d.lda;
#pragma unroll
      for (int e = 0; e < EA; ++e) ra[e] = src[e];
    }
    if (TB == 0) {
      const T* src = B + (k0 + b_c) + (j0 + b_r) * d.ldb;
#pragma unroll
      for (int e = 0; e < EB; ++e) rb[e] = src[e];
    } else {
      const T* src = B + (j0 + b_c) + (k0 + b_r) * d.ldb;
#pragma unroll
      for (int e = 0; e < EB; ++e) rb[e] = src[e];
    }
  };
  auto load_tiles = [&](int64_t k0) {
    if (interior && k0 + BK <= ke) {
      load_fast(k0);
      return;
    }
    if (TA == 0) {
      const int64_t k = k0 + a_r;
      const bool kin = k < ke;
#pragma unroll
      for (int e = 0; e < EA; ++e) {
        const int64_t i = i0 + a_c + e;
        ra[e] = (kin && i < d.M) ? A[i + k * d.lda] : T(0);
      }
    } else {
      const int64_t i = i0 + a_r;
#pragma unroll
      for (int e = 0; e < EA; ++e) {
        const int64_t k = k0 + a_c + e;
        ra[e] = (i < d.M && k < ke) ? A[k + i * d.lda] : T(0);
      }
    }
    if (TB == 0) {
      const int64_t j = j0 + b_r;
#pragma unroll
      for (int e = 0; e < EB; ++e) {
        const int64_t k = k0 + b_c + e;
        rb[e] = (j < d.N && k < ke) ? B[k + j * d.ldb] : T(0);
      }
    } else {
      const int64_t k = k0 + b_r;
      const bool kin = k < ke;
#pragma unroll
      for (int e = 0; e < EB; ++e) {
        const int64_t j = j0 + b_c + e;
        rb[e] = (kin && j < d.N) ? B[j + k * d.ldb] : T(0);
      }
    }
  };
  auto store_tiles = [&](int b) {
    if (TA == 0) {
#pragma unroll
      for (int e = 0; e < EA; ++e) As[b][a_r][a_c + e] = ra[e];
    } else {
#pragma unroll
      for (int e = 0; e < EA; ++e) As[b][a_c + e][a_r] = ra[e];
    }
    if (TB == 0) {
#pragma unroll
      for (int e = 0; e < EB; ++e) Bs[b][b_c + e][b_r] = rb[e];
    } else {
#pragma unroll
      for (int e = 0; e < EB; ++e) Bs[b][b_r][b_c + e] = rb[e];
    }
  };

  auto compute = [&](int buf, int ks) {
    const int kr = ks + (lane >> 4);
    T av[FM], bv[FN];
#pragma unroll
    for (int m = 0; m < FM; ++m) av[m] = As[buf][kr][wm * (FM * 16) + m * 16 + (lane & 15)];
#pragma unroll
    for (int n = 0; n < FN; ++n) bv[n] = Bs[buf][kr][wn * (FN * 16) + n * 16 + (lane & 15)];
#pragma unroll
    for (int m = 0; m < FM; ++m)
#pragma unroll
      for (int n = 0; n < FN; ++n) acc[m][n] = mfma16(bv[n], av[m], acc[m][n]);
  };

  if (kb < ke) {
    // Two LDS stages and a register stage: while slab k feeds the MFMAs from
    // one LDS buffer, slab k+1 (loaded into registers during slab k-1) is
    // written to the other buffer right after the first k-step, and the global
    // loads of slab k+2 are issued behind it.  The loads thus have a whole slab
    // of MFMAs to land, and the single barrier per slab waits for LDS traffic
    // only, not for the loads in flight.
    load_tiles(kb);
    store_tiles(0);
    if (kb + BK < ke) load_tiles(kb + BK);
    lds_barrier();
    int buf = 0;
    for (int64_t k0 = kb; k0 < ke; k0 += BK) {
      compute(buf, 0);
      if (k0 + BK < ke) {
        store_tiles(buf ^ 1);
        if (k0 + 2 * BK < ke) load_tiles(k0 + 2 * BK);
      }
#pragma unroll
      for (int ks = 4; ks < BK; ks += 4) compute(buf, ks);
      lds_barrier();
      buf ^= 1;
    }
  }

  // The MFMAs compute C^T tiles (B fragment as the row operand), so each lane
  // holds one row of C and the 16 lanes of a row group cover 16 consecutive
  // rows of a column: stores and the beta loads are 128-byte (FP64) or
  // 64-byte (FP32) contiguous segments.  Lane layout of the 16x16x4 outputs:
  // column 4*v + lane/16 (f64) or 4*(lane/16) + v (f32), row lane%16.
  const double alpha = d.alphaVec ? d.alpha * d.alphaVec[z] : d.alpha;
  const T al = T(alpha), be = T(d.beta);
#pragma unroll
  for (int m = 0; m < FM; ++m) {
    const int64_t row = i0 + wm * (FM * 16) + m * 16 + (lane & 15);
    if (row >= d.M) continue;
#pragma unroll
    for (int n = 0; n < FN; ++n) {
      const int64_t colb = j0 + wn * (FN * 16) + n * 16;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t col = colb + ((sizeof(T) == 8) ? (4 * v + (lane >> 4)) : ((lane >> 4) * 4 + v));
        if (col < d.N) {
          T* p = C + row + col * d.ldc;
          const T val = al * acc[m][n][v];
          *p = (d.beta == 0.0) ? val : val + be * (*p);
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// BF16x3 kernel: fp32 in / fp32 out, three bf16 MFMA products per k-step.
// ---------------------------------------------------------------------------
__device__ inline unsigned short bf16_rn(float x) {
  unsigned int u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ inline float bf16_to_f(unsigned short h) { return __uint_as_float(((unsigned int)h) << 16); }

template <int TA, int TB>
__global__ __launch_bounds__(kThreads) void gemm_bf16x3_kernel(GemmDesc d) {
  constexpr int BK = 32;
  constexpr int LDK = BK + 8;  // row pitch in bf16 elements (80 bytes)
  __shared__ unsigned short Ah[BM][LDK], Al[BM][LDK], Bh[BN][LDK], Bl[BN][LDK];

  const TileCoord tc = tile_of<BM, BN>(d.M, d.N, d.lowerC);
  if (tc.skip) return;
  const int64_t i0 = tc.i0, j0 = tc.j0;
  const int64_t z = blockIdx.z;
  const float* __restrict__ A = reinterpret_cast<const float*>(d.A) + (d.mapA ? (int64_t)d.mapA[z] : z) * d.strideA;
  const float* __restrict__ B = reinterpret_cast<const float*>(d.B) + (d.mapB ? (int64_t)d.mapB[z] : z) * d.strideB;
  float* __restrict__ C = reinterpret_cast<float*>(d.C) + (d.mapC ? (int64_t)d.mapC[z] : z) * d.strideC;

  int64_t kb, ke;
  k_range<BM, BN>(d, i0, j0, kb, ke);

  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int wm = wave >> 1, wn = wave & 1;

  f4 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = f4{0, 0, 0, 0};

  float ra[16], rb[16];
  const bool interior = (i0 + BM <= d.M) && (j0 + BN <= d.N);
  auto load_fast = [&](int64_t k0) {
    if (TA == 0) {
      const float* src = A + (i0 + (t & 7) * 16) + (k0 + (t >> 3)) * d.lda;
#pragma unroll
      for (int e = 0; e < 16; ++e) ra[e] = src[e];
    } else {
      const float* src = A + (k0 + (t & 1) * 16) + (i0 + (t >> 1)) * d.lda;
#pragma unroll
      for (int e = 0; e < 16; ++e) ra[e] = src[e];
    }
    if (TB == 0) {
      const float* src = B + (k0 + (t & 1) * 16) + (j0 + (t >> 1)) * d.ldb;
#pragma unroll
      for (int e = 0; e < 16; ++e) rb[e] = src[e];
    } else {
      const float* src = B + (j0 + (t & 7) * 16) + (k0 + (t >> 3)) * d.ldb;
#pragma unroll
      for (int e = 0; e < 16; ++e) rb[e] = src[e];
    }
  };
  auto load_tiles = [&](int64_t k0) {
    if (interior && k0 + BK <= ke) {
      load_fast(k0);
      return;
    }
    if (TA == 0) {  // thread: k = t>>3, 16 consecutive i
      const int kk = t >> 3, ii = (t & 7) * 16;
      const int64_t k = k0 + kk;
      const bool kin = k < ke;
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int64_t i = i0 + ii + e;
        ra[e] = (kin && i < d.M) ? A[i + k * d.lda] : 0.f;
      }
    } else {  // thread: i = t>>1, 16 consecutive k
      const int ii = t >> 1, kk = (t & 1) * 16;
      const int64_t i = i0 + ii;
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int64_t k = k0 + kk + e;
        ra[e] = (i < d.M && k < ke) ? A[k + i * d.lda] : 0.f;
      }
    }
    if (TB == 0) {  // thread: j = t>>1, 16 consecutive k
      const int jj = t >> 1, kk = (t & 1) * 16;
      const int64_t j = j0 + jj;
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int64_t k = k0 + kk + e;
        rb[e] = (j < d.N && k < ke) ? B[k + j * d.ldb] : 0.f;
      }
    } else {  // thread: k = t>>3, 16 consecutive j
      const int kk = t >> 3, jj = (t & 7) * 16;
      const int64_t k = k0 + kk;
      const bool kin = k < ke;
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int64_t j = j0 + jj + e;
        rb[e] = (kin && j < d.N) ? B[j + k * d.ldb] : 0.f;
      }
    }
  };
  auto split_store = [&](unsigned short (*H)[LDK], unsigned short (*L)[LDK], int r, int c, float x) {
    const unsigned short h = bf16_rn(x);
    H[r][c] = h;
    L[r][c] = bf16_rn(x - bf16_to_f(h));
  };
  auto store_tiles = [&]() {
    if (TA == 0) {
      const int kk = t >> 3, ii = (t & 7) * 16;
#pragma unroll
      for (int e = 0; e < 16; ++e) split_store(Ah, Al, ii + e, kk, ra[e]);
    } else {
      const int ii = t >> 1, kk = (t & 1) * 16;
#pragma unroll
      for (int e = 0; e < 16; ++e) split_store(Ah, Al, ii, kk + e, ra[e]);
    }
    if (TB == 0) {
      const int jj = t >> 1, kk = (t & 1) * 16;
#pragma unroll
      for (int e = 0; e < 16; ++e) split_store(Bh, Bl, jj, kk + e, rb[e]);
    } else {
      const int kk = t >> 3, jj = (t & 7) * 16;
#pragma unroll
      for (int e = 0; e < 16; ++e) split_store(Bh, Bl, jj + e, kk, rb[e]);
    }
  };

  if (kb < ke) {
    load_tiles(kb);
    store_tiles();
    __syncthreads();
    for (int64_t k0 = kb; k0 < ke; k0 += BK) {
      const bool more = (k0 + BK) < ke;
      if (more) load_tiles(k0 + BK);
      const int kc = (lane >> 4) * 8;
      bf8 ah[4], al[4], bh[4], bl[4];
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        const int r = wm * 64 + m * 16 + (lane & 15);
        ah[m] = __builtin_bit_cast(bf8, *reinterpret_cast<const u16x8*>(&Ah[r][kc]));
        al[m] = __builtin_bit_cast(bf8, *reinterpret_cast<const u16x8*>(&Al[r][kc]));
      }
#pragma unroll
      for (int n = 0; n < 4; ++n) {
        const int r = wn * 64 + n * 16 + (lane & 15);
        bh[n] = __builtin_bit_cast(bf8, *reinterpret_cast<const u16x8*>(&Bh[r][kc]));
        bl[n] = __builtin_bit_cast(bf8, *reinterpret_cast<const u16x8*>(&Bl[r][kc]));
      }
#pragma unroll
      for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int n = 0; n < 4; ++n) {
          acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[m], bh[n], acc[m][n], 0, 0, 0);
          acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[m], bl[n], acc[m][n], 0, 0, 0);
          acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[m], bh[n], acc[m][n], 0, 0, 0);
        }
      __syncthreads();
      if (more) {
        store_tiles();
        __syncthreads();
      }
    }
  }

  const double alpha = d.alphaVec ? d.alpha * d.alphaVec[z] : d.alpha;
  const float alf = float(alpha), be = float(d.beta);
#pragma unroll
  for (int m = 0; m < 4; ++m) {
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      const int64_t col = j0 + wn * 64 + n * 16 + (lane & 15);
      if (col >= d.N) continue;
      const int64_t rowb = i0 + wm * 64 + m * 16 + (lane >> 4) * 4;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t row = rowb + v;
        if (row < d.M) {
          float* p = C + row + col * d.ldc;
          const float val = alf * acc[m][n][v];
          *p = (d.beta == 0.0) ? val : val + be * (*p);
        }
      }
    }
  }
}

// C = alpha_z * sum_y P[z][y] + beta * C  (split-K reduction, fixed order)
template <typename T>
__global__ __launch_bounds__(kThreads) void splitk_reduce_kernel(GemmDesc d, const T* P) {
  const int64_t z = blockIdx.y;
  const int64_t mn = d.M * d.N;
  const double alpha = d.alphaVec ? d.alpha * d.alphaVec[z] : d.alpha;
  T* C = reinterpret_cast<T*>(d.C) + (d.mapC ? (int64_t)d.mapC[z] : z) * d.strideC;
  const T* Pz = P + z * d.splitk * mn;
  for (int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x; e < mn; e += (int64_t)gridDim.x * kThreads) {
    const int64_t i = e % d.M, j = e / d.M;
    double s = 0.0;
    for (int y = 0; y < d.splitk; ++y) s += (double)Pz[y * mn + e];
    T* p = C + i + j * d.ldc;
    const double v = alpha * s;
    *p = (d.beta == 0.0) ? (T)v : (T)(v + d.beta * (double)(*p));
  }
}

bool g_capture = false;

void* g_splitk_ws = nullptr;
size_t g_splitk_cap = 0;

int splitk_workspace(size_t bytes, void** p) {
  if (bytes > g_splitk_cap) {
    if (g_splitk_ws) {
      ZOLO_HIP(hipDeviceSynchronize());
      ZOLO_HIP(hipFree(g_splitk_ws));
    }
    g_splitk_ws = nullptr;
    g_splitk_cap = 0;
    ZOLO_HIP(hipMalloc(&g_splitk_ws, bytes));
    g_splitk_cap = bytes;
  }
  *p = g_splitk_ws;
  return kOk;
}

struct LedgerScope {
  bool on;
  LedgerEntry e;
  hipStream_t s;
  LedgerScope(const GemmDesc& d, int prec, int tm, int tn, hipStream_t st) : on(g_ledger_on && !g_capture), s(st) {
    if (!on) return;
    e.prec = prec; e.ta = d.transA; e.tb = d.transB; e.tm = tm; e.tn = tn;
    e.flops = algorithmic_flops(d);
    if (hipEventCreate(&e.e0) != hipSuccess || hipEventCreate(&e.e1) != hipSuccess) { on = false; return; }
    (void)hipEventRecord(e.e0, s);
  }
  ~LedgerScope() {
    if (!on) return;
    (void)hipEventRecord(e.e1, s);
    g_ledger.push_back(e);
  }
};

template <typename T, int TBM, int TBN>
int launch_tile(const GemmDesc& d0, hipStream_t s) {
  const int64_t tiles = cdiv(d0.M, TBM) * cdiv(d0.N, TBN);
  LedgerScope ls(d0, sizeof(T) == 8 ? 0 : 1, TBM, TBN, s);
  // Few output tiles over a long K (the width-64 V^T C products of the QR):
  // split K so the chip has enough workgroups, then reduce in a fixed order.
  GemmDesc d = d0;
  int split = 1;
  const int64_t wgs = tiles * d0.batch;
  if (!g_capture && !d0.lowerC && wgs < 512 && d0.K >= 1024) {
    split = (int)std::min<int64_t>(std::min<int64_t>(32, cdiv(1024, wgs)), d0.K / 512);
    if (split < 2) split = 1;
  }
  if (split > 1) {
    void* ws = nullptr;
    ZOLO_CHECK(splitk_workspace(sizeof(T) * (size_t)split * d0.batch * d0.M * d0.N, &ws));
    d.splitk = split;
    d.kchunk = cdiv(cdiv(d0.K, split), 32) * 32;
    d.splitStride = d0.M * d0.N;
    d.C = ws;
    d.ldc = d0.M;
    d.mapC = nullptr;
    d.strideC = (int64_t)split * d0.M * d0.N;
    d.alpha = 1.0;
    d.alphaVec = nullptr;
    d.beta = 0.0;
  }
  dim3 grid((unsigned)tiles, (unsigned)split, (unsigned)d.batch);
  if (d.transA == 0 && d.transB == 0) hipLaunchKernelGGL((gemm_kernel<T, 0, 0, TBM, TBN>), grid, dim3(kThreads), 0, s, d);
  else if (d.transA == 0) hipLaunchKernelGGL((gemm_kernel<T, 0, 1, TBM, TBN>), grid, dim3(kThreads), 0, s, d);
  else if (d.transB == 0) hipLaunchKernelGGL((gemm_kernel<T, 1, 0, TBM, TBN>), grid, dim3(kThreads), 0, s, d);
  else hipLaunchKernelGGL((gemm_kernel<T, 1, 1, TBM, TBN>), grid, dim3(kThreads), 0, s, d);
  ZOLO_HIP(hipGetLastError());
  if (split > 1) {
    GemmDesc r = d0;
    r.splitk = split;
    const int64_t mn = d0.M * d0.N;
    const unsigned gx = (unsigned)std::min<int64_t>(4096, std::max<int64_t>(1, cdiv(mn, kThreads)));
    hipLaunchKernelGGL((splitk_reduce_kernel<T>), dim3(gx, d0.batch), dim3(kThreads), 0, s, r,
                       reinterpret_cast<const T*>(d.C));
    ZOLO_HIP(hipGetLastError());
  }
  return kOk;
}

// Skinny operands (the width-64 compact-WY factors of the QR) take a 64-wide
// tile so no MFMA work is spent on padding.
template <typename T>
int launch_typed(const GemmDesc& d, hipStream_t s) {
  switch (d.tile) {
    case 1: return launch_tile<T, 128, 128>(d, s);
    case 2: return launch_tile<T, 64, 128>(d, s);
    case 3: return launch_tile<T, 128, 64>(d, s);
    case 4: return launch_tile<T, 64, 64>(d, s);
    default: break;
  }
  if (d.M <= 64 && d.N <= 64) return launch_tile<T, 64, 64>(d, s);
  if (d.M <= 64 && d.lowerC == 0) return launch_tile<T, 64, 128>(d, s);
  if (d.N <= 64 && d.lowerC == 0) return launch_tile<T, 128, 64>(d, s);
  return launch_tile<T, 128, 128>(d, s);
}

int launch_bf16x3(const GemmDesc& d, hipStream_t s) {
  const int64_t tiles = cdiv(d.M, BM) * cdiv(d.N, BN);
  dim3 grid((unsigned)tiles, 1, (unsigned)d.batch);
  LedgerScope ls(d, 2, BM, BN, s);
  if (d.transA == 0 && d.transB == 0) hipLaunchKernelGGL((gemm_bf16x3_kernel<0, 0>), grid, dim3(kThreads), 0, s, d);
  else if (d.transA == 0) hipLaunchKernelGGL((gemm_bf16x3_kernel<0, 1>), grid, dim3(kThreads), 0, s, d);
  else if (d.transB == 0) hipLaunchKernelGGL((gemm_bf16x3_kernel<1, 0>), grid, dim3(kThreads), 0, s, d);
  else hipLaunchKernelGGL((gemm_bf16x3_kernel<1, 1>), grid, dim3(kThreads), 0, s, d);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

}  // namespace

void gemm_set_capture(bool on) { g_capture = on; }

int gemm(Prec p, const GemmDesc& d, hipStream_t s) {
  if (d.M <= 0 || d.N <= 0 || d.batch <= 0) return kOk;
  if (d.K <= 0 && d.beta == 1.0) return kOk;
  if (cdiv(d.M, BM) * cdiv(d.N, BN) > 0x7fffffffLL || d.batch > 65535) {
    set_error("gemm: grid too large");
    return kErrArg;
  }
  switch (p) {
    case Prec::F64: return launch_typed<double>(d, s);
    case Prec::F32: return launch_typed<float>(d, s);
    case Prec::BF16X3: return launch_bf16x3(d, s);
  }
  return kErrArg;
}

}  // namespace zolo
