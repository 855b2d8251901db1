// LDS-tiled MFMA GEMM engine for gfx950 (CDNA4), column-major operands.
//
// One engine serves every dense contraction of the polar-decomposition path:
// the Gram matrix X^T X (lower triangle only), the Cholesky trailing updates,
// the triangular multiplies by explicit inverse factors, the compact-WY
// updates of the structured Householder QR and the Q1 Q2^T product.
//
//  * F64    : v_mfma_f64_16x16x4f64, 128x128x16 tile, 4 waves of 64x64.
//  * F32    : v_mfma_f32_16x16x4f32, same tiling.
//  * BF16X3 : fp32 operands split once per call into hi + lo bfloat16 halves
//             (a tiled layout whose K slabs are contiguous), three
//             v_mfma_f32_16x16x32_bf16 products (lo*hi + hi*lo + hi*hi)
//             accumulated in fp32; 256x256x32 tile of 8 waves of 128x64.
//             Small products (M N K batch <= kB3MinWork, the deep levels of the
//             recursive Cholesky and inverse) run on the F32 path instead:
//             the two split launches and the 256x256 tile padding would cost
//             more than the FP32 MFMA rate, and FP32 products are the more
//             accurate of the two.
//
// Work distribution: tiles are numbered so that the 8 XCDs (round-robin
// workgroup dispatch) each own a contiguous block of tiles, and inside a block
// tiles walk down groups of 8 tile-rows so neighbouring workgroups share the
// same A and B panels in their XCD's L2.
#include <algorithm>
#include <vector>

#include "common.hpp"
#include "gemm_tiles.hpp"

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

// Number of (i, j, k) with i < M, j < N and max(0, i - alo, j - blo) <= k < K:
// the products of op(A) zero for k < i - alo times op(B) zero for k < j - blo.
// For each i the count over j is two arithmetic series.
double two_band_triples(int64_t M, int64_t N, int64_t K, int64_t alo, int64_t blo) {
  double t = 0.0;
  for (int64_t i = 0; i < M; ++i) {
    const int64_t c = std::max<int64_t>(0, i - alo);
    if (c >= K) continue;
    // j <= c + blo: k from c; j > c + blo: k from j - blo (until K)
    const int64_t j1 = std::min<int64_t>(N, c + blo + 1);
    t += (double)j1 * (double)(K - c);
    const int64_t j2 = std::min<int64_t>(N, K + blo);  // j - blo < K
    if (j2 > j1) {
      const double cnt = (double)(j2 - j1);
      const double first = (double)(K - (j1 - blo)), last = (double)(K - (j2 - 1 - blo));
      t += cnt * (first + last) / 2.0;
    }
  }
  return t;
}

double algorithmic_flops(const GemmDesc& d) {
  // 2 M N K scaled by the nonzero fraction of each banded operand (exact for
  // the triangular and trapezoidal operands of the path) and by the computed
  // fraction of C for lower-triangle-only outputs; both operands with lower
  // bands: the exact count of contributing products.
  if (d.aLo >= 0 && d.bLo >= 0 && d.aHi < 0 && d.bHi < 0 && !d.lowerC && d.K > 0)
    return 2.0 * two_band_triples(d.M, d.N, d.K, d.aLo, d.bLo) * d.batch;
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
constexpr int BM = 128, BN = 128;   // smallest tile of the grid-size check

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

// ---------------------------------------------------------------------------
// F64 / F32 kernel.  Tile TBM x TBN x 16, LDS double-buffered; the four waves form a WM x WN grid,
// each wave owning FM x FN blocks of 16x16 MFMA outputs.
// ---------------------------------------------------------------------------
template <typename T, int TA, int TB, int TBM, int TBN>
__global__ __launch_bounds__(kThreads, 2) void gemm_kernel(GemmDesc d) {
  constexpr int BK = 16;
  constexpr int PAD = (sizeof(T) == 8) ? 4 : 8;
  constexpr int WM = (TBM == TBN) ? 2 : (TBM == 64 ? 1 : 4);
  constexpr int WN = 4 / WM;
  constexpr int FM = TBM / WM / 16, FN = TBN / WN / 16;
  constexpr int EA = TBM * BK / kThreads;   // A elements per thread
  constexpr int EB = TBN * BK / kThreads;   // B elements per thread
  __shared__ T As[2][BK][TBM + PAD];
  __shared__ T Bs[2][BK][TBN + PAD];

  const TileCoord tc = tile_of<TBM, TBN>(d.M, d.N, d.lowerC, banded(d), heavy_right(d));
  if (tc.skip) return;
  const int64_t i0 = tc.i0, j0 = tc.j0;
  const int64_t z = blockIdx.z;
  const T* __restrict__ A = reinterpret_cast<const T*>(d.A) + (d.mapA ? (int64_t)d.mapA[z] : z) * d.strideA;
  const T* __restrict__ B = reinterpret_cast<const T*>(d.B) + (d.mapB ? (int64_t)d.mapB[z] : z) * d.strideB;
  T* __restrict__ C = reinterpret_cast<T*>(d.C) + (d.mapC ? (int64_t)d.mapC[z] : z) * d.strideC +
                      ((d.cGapAt >= 0 && i0 >= d.cGapAt) ? d.cGap : 0);

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
  // the second row range of B (transB = 0): a slab lies on one side of bGapAt
  auto boff = [&](int64_t k0) -> int64_t { return (d.bGapAt >= 0 && k0 >= d.bGapAt) ? d.bGap : 0; };
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
      const T* src = B + (k0 + b_c + boff(k0)) + (j0 + b_r) * d.ldb;
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
        rb[e] = (j < d.N && k < ke) ? B[k + boff(k0) + j * d.ldb] : T(0);
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
  // With beta != 0 every row group reads all its C values before it writes
  // any: the loads go out together and are waited for once (a load after a
  // store to C could alias it, so the compiler would otherwise wait for each
  // element in turn).
  const double alpha = d.alphaVec ? d.alpha * d.alphaVec[z] : d.alpha;
  const T al = T(alpha), be = T(d.beta);
  const bool rd = d.beta != 0.0;
  if (interior) {  // no bounds checks: straight-line loads and stores
#pragma unroll
    for (int m = 0; m < FM; ++m) {
      T* Cr = C + i0 + wm * (FM * 16) + m * 16 + (lane & 15);
      T cv[FN][4];
      if (rd) {
#pragma unroll
        for (int n = 0; n < FN; ++n)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int64_t col = j0 + wn * (FN * 16) + n * 16 +
                                ((sizeof(T) == 8) ? (4 * v + (lane >> 4)) : ((lane >> 4) * 4 + v));
            cv[n][v] = Cr[col * d.ldc];
          }
      }
#pragma unroll
      for (int n = 0; n < FN; ++n)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int64_t col = j0 + wn * (FN * 16) + n * 16 +
                              ((sizeof(T) == 8) ? (4 * v + (lane >> 4)) : ((lane >> 4) * 4 + v));
          const T val = al * acc[m][n][v];
          Cr[col * d.ldc] = rd ? val + be * cv[n][v] : val;
        }
    }
    return;
  }
#pragma unroll
  for (int m = 0; m < FM; ++m) {
    const int64_t row = i0 + wm * (FM * 16) + m * 16 + (lane & 15);
    if (row >= d.M) continue;
    T cv[FN][4];
#pragma unroll
    for (int n = 0; n < FN; ++n) {
      const int64_t colb = j0 + wn * (FN * 16) + n * 16;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t col = colb + ((sizeof(T) == 8) ? (4 * v + (lane >> 4)) : ((lane >> 4) * 4 + v));
        cv[n][v] = (rd && col < d.N) ? C[row + col * d.ldc] : T(0);
      }
    }
#pragma unroll
    for (int n = 0; n < FN; ++n) {
      const int64_t colb = j0 + wn * (FN * 16) + n * 16;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t col = colb + ((sizeof(T) == 8) ? (4 * v + (lane >> 4)) : ((lane >> 4) * 4 + v));
        if (col < d.N) {
          const T val = al * acc[m][n][v];
          C[row + col * d.ldc] = rd ? val + be * cv[n][v] : val;
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

constexpr int kSplitTile = 64;
constexpr int kB3K = 32;            // K depth of one bf16x3 slab and of a layout tile
constexpr int kB3Rows = 128;        // rows of a layout tile
constexpr int kB3M = 256, kB3N = 256, kB3Threads = 512;  // workgroup tile: 8 waves of 128 x 64
constexpr double kB3MinWork = 2147483648.0;  // M N K batch at or below which BF16X3 takes the F32 path

// Offset of element (r, k) of a split operand in its tiled layout: tiles of
// 128 rows x 32 k (8 KB of bf16, row-major inside), ordered k-fastest within a
// 128-row block, so the K slabs of one row block are consecutive in memory.
__device__ inline int64_t b3_offset(int64_t r, int64_t k, int64_t Kp) {
  return ((r / kB3Rows) * (Kp / kB3K) + k / kB3K) * (kB3Rows * kB3K) + (r % kB3Rows) * kB3K + (k % kB3K);
}

// Split an fp32 operand into bf16 hi and lo halves (x = hi + lo to 16 bits) in
// the tiled layout, zero outside R x K (rows padded to Rp, depth to Kp).
// Source element (r, k) is src[k + r * ld] when KMAJ, else src[r + k * ld].
// Batch item z reads operand map ? map[z] : z (stride 0: one shared operand).
template <int KMAJ>
__global__ __launch_bounds__(kThreads) void bf16_split_kernel(const float* __restrict__ src, int64_t ld,
                                                              int64_t stride, const int* map, int64_t R, int64_t K,
                                                              int64_t Kp, unsigned short* __restrict__ hi,
                                                              unsigned short* __restrict__ lo, int64_t ostride) {
  __shared__ float tile[kSplitTile][kSplitTile + 1];
  const int64_t z = blockIdx.z;
  const float* S = src + (map ? (int64_t)map[z] : z) * stride;
  const int64_t k0 = (int64_t)blockIdx.x * kSplitTile, r0 = (int64_t)blockIdx.y * kSplitTile;
  const int t = threadIdx.x;
#pragma unroll
  for (int it = 0; it < kSplitTile * kSplitTile / kThreads; ++it) {
    const int idx = t + it * kThreads;
    const int rr = KMAJ ? idx / kSplitTile : idx % kSplitTile;
    const int kk = KMAJ ? idx % kSplitTile : idx / kSplitTile;
    const int64_t r = r0 + rr, k = k0 + kk;
    float v = 0.f;
    if (r < R && k < K) v = KMAJ ? S[k + r * ld] : S[r + k * ld];
    tile[rr][kk] = v;
  }
  __syncthreads();
  // thread: one row, 16 consecutive k (inside one 32-wide layout tile) -> two 16-byte stores per half
  const int rr = t >> 2, kk = (t & 3) * 16;
  unsigned int ph[8], pl[8];
#pragma unroll
  for (int e = 0; e < 16; e += 2) {
    const float x0 = tile[rr][kk + e], x1 = tile[rr][kk + e + 1];
    const unsigned short h0 = bf16_rn(x0), h1 = bf16_rn(x1);
    const unsigned short l0 = bf16_rn(x0 - bf16_to_f(h0)), l1 = bf16_rn(x1 - bf16_to_f(h1));
    ph[e / 2] = (unsigned int)h0 | ((unsigned int)h1 << 16);
    pl[e / 2] = (unsigned int)l0 | ((unsigned int)l1 << 16);
  }
  const int64_t o = z * ostride + b3_offset(r0 + rr, k0 + kk, Kp);
  uint4* H = reinterpret_cast<uint4*>(hi + o);
  uint4* L = reinterpret_cast<uint4*>(lo + o);
  H[0] = make_uint4(ph[0], ph[1], ph[2], ph[3]);
  H[1] = make_uint4(ph[4], ph[5], ph[6], ph[7]);
  L[0] = make_uint4(pl[0], pl[1], pl[2], pl[3]);
  L[1] = make_uint4(pl[4], pl[5], pl[6], pl[7]);
}

// BF16x3 GEMM on pre-split operands in the tiled layout: op(A) rows and op(B)
// columns (hi/lo, both padded to 256-multiples).  256 x 256 workgroup tile of
// 8 waves, each 128 x 64 (8 x 4 blocks of 16 x 16), so that a slab's 96
// MFMAs per wave amortise the slab's barrier and LDS latency; K slabs of 32
// double-buffered in LDS with a register stage (slab k+2 loads in flight while
// slab k computes), one LDS-only barrier per slab.  Per 16x16 block: three
// v_mfma_f32_16x16x32_bf16 (lo*hi + hi*lo + hi*hi), fp32 accumulation; the
// products with hi(A) go first, then those with lo(A), so only one half of the
// A fragments is live at a time.
__global__ __launch_bounds__(kB3Threads) void gemm_bf16x3_kernel(GemmDesc d, const unsigned short* __restrict__ Ahi,
                                                                 const unsigned short* __restrict__ Alo, int64_t sAs,
                                                                 const unsigned short* __restrict__ Bhi,
                                                                 const unsigned short* __restrict__ Blo, int64_t sBs,
                                                                 int64_t Kp) {
  __shared__ __attribute__((aligned(16))) unsigned short LA[2][2][kB3M][kB3K];  // [stage][hi lo][row][k, swizzled]
  __shared__ __attribute__((aligned(16))) unsigned short LB[2][2][kB3N][kB3K];

  const TileCoord tc = tile_of<kB3M, kB3N>(d.M, d.N, d.lowerC, banded(d), heavy_right(d));
  if (tc.skip) return;
  const int64_t i0 = tc.i0, j0 = tc.j0;
  const int64_t z = blockIdx.z;
  float* __restrict__ C = reinterpret_cast<float*>(d.C) + (d.mapC ? (int64_t)d.mapC[z] : z) * d.strideC;
  const unsigned short* __restrict__ Ah = Ahi + z * sAs;
  const unsigned short* __restrict__ Al = Alo + z * sAs;
  const unsigned short* __restrict__ Bh = Bhi + z * sBs;
  const unsigned short* __restrict__ Bl = Blo + z * sBs;

  int64_t kb, ke;
  k_range<kB3M, kB3N>(d, i0, j0, kb, ke);
  kb = (kb / kB3K) * kB3K;
  ke = min((ke + kB3K - 1) / kB3K * kB3K, Kp);

  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int wm = wave >> 2, wn = wave & 3;  // wave tile: rows wm*128 .. +128, columns wn*64 .. +64

  f4 acc[8][4];
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = f4{0, 0, 0, 0};

  // A and B slabs: 2 layout tiles each (512 chunks of 16 B per tile and
  // half); thread t takes chunk t of each.  Chunk e of a tile is row e / 4,
  // k (e % 4) * 8.  The register stage as named vectors (an array of them was
  // kept on the scratch stack by the compiler).
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  u32x4 ah0, ah1, al0, al1, bh0, bh1, bl0, bl1;
  const int64_t tiles_k = Kp / kB3K;
  const int64_t tile_elems = tiles_k * (kB3Rows * kB3K);  // one 128-row block over all of K
  const int64_t oa0 = (i0 / kB3Rows) * tile_elems + (int64_t)t * 8;
  const int64_t ob0 = (j0 / kB3Rows) * tile_elems + (int64_t)t * 8;
  auto load_tiles = [&](int64_t k0) {
    const int64_t ko = (k0 / kB3K) * (kB3Rows * kB3K);
    ah0 = *reinterpret_cast<const u32x4*>(Ah + oa0 + ko);
    al0 = *reinterpret_cast<const u32x4*>(Al + oa0 + ko);
    ah1 = *reinterpret_cast<const u32x4*>(Ah + oa0 + tile_elems + ko);
    al1 = *reinterpret_cast<const u32x4*>(Al + oa0 + tile_elems + ko);
    bh0 = *reinterpret_cast<const u32x4*>(Bh + ob0 + ko);
    bl0 = *reinterpret_cast<const u32x4*>(Bl + ob0 + ko);
    bh1 = *reinterpret_cast<const u32x4*>(Bh + ob0 + tile_elems + ko);
    bl1 = *reinterpret_cast<const u32x4*>(Bl + ob0 + tile_elems + ko);
  };
  const int lrow = t >> 2, lk = lds_chunk(lrow, t & 3) * 8;  // rows lrow and 128 + lrow swizzle alike
  auto store_tiles = [&](int b) {
    *reinterpret_cast<u32x4*>(&LA[b][0][lrow][lk]) = ah0;
    *reinterpret_cast<u32x4*>(&LA[b][1][lrow][lk]) = al0;
    *reinterpret_cast<u32x4*>(&LA[b][0][kB3Rows + lrow][lk]) = ah1;
    *reinterpret_cast<u32x4*>(&LA[b][1][kB3Rows + lrow][lk]) = al1;
    *reinterpret_cast<u32x4*>(&LB[b][0][lrow][lk]) = bh0;
    *reinterpret_cast<u32x4*>(&LB[b][1][lrow][lk]) = bl0;
    *reinterpret_cast<u32x4*>(&LB[b][0][kB3Rows + lrow][lk]) = bh1;
    *reinterpret_cast<u32x4*>(&LB[b][1][kB3Rows + lrow][lk]) = bl1;
  };
  // One slab from LDS buffer b: the fragments of B and hi(A), the 64 hi(A)
  // products with the LDS stores of the staged slab k+1 (st) and the global
  // loads of slab k+2 (ld) spread among them, then lo(A) and its 32
  // products.  Spreading the stores keeps the other wave of the SIMD issuing
  // MFMAs while one waits for its staged loads (the stores bunched after the
  // MFMAs cost about 10 % of the kernel's time).
  int64_t knext = 0;
  auto slab = [&](int b, bool st, bool ld) {
    // row bits 1..2 come from lane & 15 alone (the block offsets are multiples of 16)
    const int kc = lds_chunk(lane & 15, lane >> 4) * 8;
    bf8 bh[4], bl[4], a8[8];
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      const int r = wn * 64 + n * 16 + (lane & 15);
      bh[n] = __builtin_bit_cast(bf8, *reinterpret_cast<const u16x8*>(&LB[b][0][r][kc]));
      bl[n] = __builtin_bit_cast(bf8, *reinterpret_cast<const u16x8*>(&LB[b][1][r][kc]));
    }
#pragma unroll
    for (int m = 0; m < 8; ++m)
      a8[m] = __builtin_bit_cast(bf8, *reinterpret_cast<const u16x8*>(&LA[b][0][wm * 128 + m * 16 + (lane & 15)][kc]));
    // B fragment as the row operand: the accumulators hold C^T blocks, so the
    // epilogue writes contiguous column segments
#pragma unroll
    for (int m = 0; m < 8; ++m)
#pragma unroll
      for (int n = 0; n < 4; ++n) {
        acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[n], a8[m], acc[m][n], 0, 0, 0);
        acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[n], a8[m], acc[m][n], 0, 0, 0);
      }
    if (st) store_tiles(b ^ 1);
    if (ld) load_tiles(knext);
#pragma unroll
    for (int m = 0; m < 8; ++m)
      a8[m] = __builtin_bit_cast(bf8, *reinterpret_cast<const u16x8*>(&LA[b][1][wm * 128 + m * 16 + (lane & 15)][kc]));
#pragma unroll
    for (int m = 0; m < 8; ++m)
#pragma unroll
      for (int n = 0; n < 4; ++n) acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[n], a8[m], acc[m][n], 0, 0, 0);
    // instruction order: 16 fragment reads; 8 x (8 MFMAs, 1 LDS store, 1
    // global load); the 8 lo(A) reads; the last 32 MFMAs
    __builtin_amdgcn_sched_group_barrier(0x100, 16, 0);  // LDS reads
#pragma unroll
    for (int g = 0; g < 8; ++g) {
      __builtin_amdgcn_sched_group_barrier(0x008, 8, 0);  // MFMA
      if (st) __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);  // LDS write
      if (ld) __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);  // global load
    }
    __builtin_amdgcn_sched_group_barrier(0x100, 8, 0);
    __builtin_amdgcn_sched_group_barrier(0x008, 32, 0);
  };

  if (kb < ke) {
    load_tiles(kb);
    store_tiles(0);
    if (kb + kB3K < ke) load_tiles(kb + kB3K);
    lds_barrier();
    int buf = 0;
    int64_t k0 = kb;
    // the loop body is straight-line code so the schedule above holds in it
    for (; k0 + 2 * kB3K < ke; k0 += kB3K) {
      knext = k0 + 2 * kB3K;
      slab(buf, true, true);
      lds_barrier();
      buf ^= 1;
    }
    if (k0 + kB3K < ke) {
      slab(buf, true, false);
      lds_barrier();
      buf ^= 1;
    }
    slab(buf, false, false);
  }

  // lane layout of the 16x16x32 outputs (C^T blocks): row lane%16, column 4*(lane/16) + v
  // (beta != 0: each row group's C values are all read before any is written)
  const double alpha = d.alphaVec ? d.alpha * d.alphaVec[z] : d.alpha;
  const float alf = float(alpha), be = float(d.beta);
  const bool rd = d.beta != 0.0;
  if (i0 + kB3M <= d.M && j0 + kB3N <= d.N) {  // interior tile: straight-line loads and stores
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      float* Cr = C + i0 + wm * 128 + m * 16 + (lane & 15);
      float cv[4][4];
      if (rd) {
#pragma unroll
        for (int n = 0; n < 4; ++n)
#pragma unroll
          for (int v = 0; v < 4; ++v) cv[n][v] = Cr[(j0 + wn * 64 + n * 16 + (lane >> 4) * 4 + v) * d.ldc];
      }
#pragma unroll
      for (int n = 0; n < 4; ++n)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const float val = alf * acc[m][n][v];
          Cr[(j0 + wn * 64 + n * 16 + (lane >> 4) * 4 + v) * d.ldc] = rd ? val + be * cv[n][v] : val;
        }
    }
    return;
  }
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int64_t row = i0 + wm * 128 + m * 16 + (lane & 15);
    if (row >= d.M) continue;
    float cv[4][4];
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      const int64_t colb = j0 + wn * 64 + n * 16 + (lane >> 4) * 4;
#pragma unroll
      for (int v = 0; v < 4; ++v) cv[n][v] = (rd && colb + v < d.N) ? C[row + (colb + v) * d.ldc] : 0.f;
    }
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      const int64_t colb = j0 + wn * 64 + n * 16 + (lane >> 4) * 4;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t col = colb + v;
        if (col < d.N) {
          const float val = alf * acc[m][n][v];
          C[row + col * d.ldc] = rd ? val + be * cv[n][v] : val;
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
  if (!g_capture && !d0.lowerC && d0.cGapAt < 0 && wgs < 512 && d0.K >= 1024) {
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

// Workspace of the split operands, grown on demand (never while capturing).
void* g_b3_ws = nullptr;
size_t g_b3_cap = 0;

int launch_bf16x3(const GemmDesc& d, hipStream_t s) {
  const int64_t Mp = cdiv(d.M, kB3M) * kB3M, Np = cdiv(d.N, kB3N) * kB3N, Kp = cdiv(d.K, kSplitTile) * kSplitTile;
  const int64_t nA = (d.strideA == 0 && !d.mapA) ? 1 : d.batch, nB = (d.strideB == 0 && !d.mapB) ? 1 : d.batch;
  const int64_t sizeA = Mp * Kp, sizeB = Np * Kp;
  const size_t bytes = sizeof(unsigned short) * 2 * (size_t)(nA * sizeA + nB * sizeB);
  if (bytes > g_b3_cap) {
    if (g_capture) {
      set_error("gemm: the bf16x3 operand workspace cannot grow during graph capture");
      return kErrArg;
    }
    if (g_b3_ws) {
      ZOLO_HIP(hipDeviceSynchronize());
      ZOLO_HIP(hipFree(g_b3_ws));
    }
    g_b3_ws = nullptr;
    g_b3_cap = 0;
    ZOLO_HIP(hipMalloc(&g_b3_ws, bytes));
    g_b3_cap = bytes;
  }
  unsigned short* Ah = reinterpret_cast<unsigned short*>(g_b3_ws);
  unsigned short* Al = Ah + nA * sizeA;
  unsigned short* Bh = Al + nA * sizeA;
  unsigned short* Bl = Bh + nB * sizeB;
  if (Kp > 0) {
    const float* A = reinterpret_cast<const float*>(d.A);
    const float* B = reinterpret_cast<const float*>(d.B);
    const dim3 ga((unsigned)(Kp / kSplitTile), (unsigned)(Mp / kSplitTile), (unsigned)nA);
    const dim3 gb((unsigned)(Kp / kSplitTile), (unsigned)(Np / kSplitTile), (unsigned)nB);
    // op(A) row i, depth k: A[k + i lda] when transposed (K-major), else A[i + k lda]
    if (d.transA)
      hipLaunchKernelGGL((bf16_split_kernel<1>), ga, dim3(kThreads), 0, s, A, d.lda, d.strideA, d.mapA, d.M, d.K, Kp,
                         Ah, Al, sizeA);
    else
      hipLaunchKernelGGL((bf16_split_kernel<0>), ga, dim3(kThreads), 0, s, A, d.lda, d.strideA, d.mapA, d.M, d.K, Kp,
                         Ah, Al, sizeA);
    // op(B) column j, depth k: B[k + j ldb] when not transposed (K-major), else B[j + k ldb]
    if (d.transB)
      hipLaunchKernelGGL((bf16_split_kernel<0>), gb, dim3(kThreads), 0, s, B, d.ldb, d.strideB, d.mapB, d.N, d.K, Kp,
                         Bh, Bl, sizeB);
    else
      hipLaunchKernelGGL((bf16_split_kernel<1>), gb, dim3(kThreads), 0, s, B, d.ldb, d.strideB, d.mapB, d.N, d.K, Kp,
                         Bh, Bl, sizeB);
    ZOLO_HIP(hipGetLastError());
  }
  const int64_t tiles = cdiv(d.M, kB3M) * cdiv(d.N, kB3N);
  dim3 grid((unsigned)tiles, 1, (unsigned)d.batch);
  LedgerScope ls(d, 2, kB3M, kB3N, s);
  hipLaunchKernelGGL(gemm_bf16x3_kernel, grid, dim3(kB3Threads), 0, s, d, Ah, Al, nA == 1 ? 0 : sizeA, Bh, Bl,
                     nB == 1 ? 0 : sizeB, Kp);
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
  if (d.bGapAt >= 0 || d.cGapAt >= 0) {
    const bool ok = (p == Prec::F64 || p == Prec::F32) &&
                    (d.bGapAt < 0 || (d.transB == 0 && d.bGapAt % 16 == 0 && d.aLo < 0 && d.bLo < 0)) &&
                    (d.cGapAt < 0 || d.cGapAt % 128 == 0) && !d.mapA && !d.mapB && !d.mapC;
    if (!ok) {
      set_error("gemm: unsupported row-range gap");
      return kErrArg;
    }
  }
  switch (p) {
    case Prec::F64: return launch_typed<double>(d, s);
    case Prec::F32: return launch_typed<float>(d, s);
    case Prec::BF16X3:
      if ((double)d.M * (double)d.N * (double)d.K * (double)d.batch <= kB3MinWork) return launch_typed<float>(d, s);
      return launch_bf16x3(d, s);
    case Prec::F64X: {
      if (!gemm_ozaki_supported(d) || d.K <= 0) return launch_typed<double>(d, s);
      LedgerScope ls(d, 3, 256, 256, s);
      return gemm_ozaki(d, s, g_capture);
    }
  }
  return kErrArg;
}

}  // namespace zolo
