// FP64 matrix products from exact int8 MFMA products (Ozaki scheme II:
// K. Ozaki, Y. Uchino, T. Imamura, "Ozaki Scheme II: A GEMM-oriented
// emulation of floating-point matrix multiplication using an integer modular
// technique", 2025).
//
// gfx950 runs v_mfma_i32_16x16x64_i8 at 64 times the rate of FP64 MFMA.  The
// product C = alpha op(A) op(B) + beta C is formed as
//   1. row scaling: op(A) row i is multiplied by 2^sA_i and rounded to
//      integers a'_ik with |a'| < 2^b (op(B) column j likewise, 2^sB_j), where
//      b is chosen so that |sum_k a'_ik b'_kj| < K 2^{2b} < M / 2 for the
//      product M of the 16 moduli below (b = 54 at K = 16384, 53 at 131072);
//   2. for each modulus m: the symmetric residues of a' and b' (int8), one
//      exact int8 GEMM with int32 accumulation (|sum| <= K 127^2 < 2^31 for
//      K <= 133000), reduced mod m (uint8 planes);
//   3. per element: the 16 residues are lifted to the unique integer in
//      (-M/2, M/2) by its mixed-radix digits (Garner's algorithm, signed
//      digits), converted to double and scaled by 2^-(sA_i + sB_j).
// The integer product is exact, so the only rounding is that of the inputs to
// b bits relative to their row (column) maximum: |error_ij| <= 2^-b (max_k
// |a_ik| sum_k |b_kj| + max_k |b_kj| sum_k |a_ik|) (+ one rounding of the
// result), as small as an FP64 GEMM's own error bound for b >= 53.
//
// Operands are stored in tiles of 128 rows x 128 k (int8, 128-byte rows) with
// the K slabs of a 128-row block consecutive.  The int8 GEMM uses the BF16x3
// engine's 256 x 256 tile of 8 waves (128 x 64 each) and a swizzled LDS
// layout of the same kind.
#include <algorithm>
#include <climits>
#include <cmath>

#include "gemm_tiles.hpp"

namespace zolo {
namespace {

constexpr int kOzMods = 16;
constexpr int kOzThreads = 256;
constexpr int kOzTile = 64;         // split-kernel tile (rows x k)
constexpr int kOzK = 128;           // K depth of a slab and of a layout tile (bytes = int8 values)
constexpr int kOzRows = 128;        // rows of a layout tile
constexpr int kOzM = 256, kOzN = 256, kOzGemmThreads = 512;

// Pairwise coprime odd moduli (greedy from 255 down): product 2^124.95.
struct OzTables {
  int m[kOzMods];
  double inv_m[kOzMods];      // 1 / m_l rounded
  int inv[kOzMods][kOzMods];  // inv[l][j] = m_j^{-1} mod m_l for j < l
  int pre[kOzMods][kOzMods];  // pre[l][j] = (m_0 ... m_{j-1}) mod m_l for j < l
  int qinv[kOzMods];          // (m_0 ... m_{l-1})^{-1} mod m_l
};
constexpr int inv_mod(int a, int m) {
  int t = 0, nt = 1, r = m, nr = a % m;
  while (nr != 0) {
    const int q = r / nr;
    const int tt = t - q * nt; t = nt; nt = tt;
    const int rr = r - q * nr; r = nr; nr = rr;
  }
  return t < 0 ? t + m : t;
}
constexpr OzTables make_tables() {
  OzTables T{{255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199, 197, 193, 191}, {}, {}, {}, {}};
  for (int l = 0; l < kOzMods; ++l) {
    T.inv_m[l] = 1.0 / T.m[l];
    for (int j = 0; j < kOzMods; ++j) T.inv[l][j] = j < l ? inv_mod(T.m[j] % T.m[l], T.m[l]) : 0;
    long long q = 1;
    for (int j = 0; j < kOzMods; ++j) {
      T.pre[l][j] = j < l ? (int)q : 0;
      if (j < l) q = q * T.m[j] % T.m[l];
    }
    T.qinv[l] = l ? inv_mod((int)q, T.m[l]) : 1;
  }
  return T;
}
constexpr OzTables kOz = make_tables();
static_assert(kOz.inv[1][0] * 255 % 253 == 1, "modular inverse table");
static_assert(kOz.inv[15][14] * 193 % 191 == 1, "modular inverse table");
static_assert(kOz.pre[2][1] == 255 % 251 && kOz.pre[3][2] == 255 * 253 % 247 && kOz.qinv[2] * (255 * 253 % 251) % 251 == 1, "mixed-radix tables");
// the mixed-radix place value of digit 8, m_0 ... m_7, split into two doubles
constexpr unsigned long long kOzP8 = 255ull * 253 * 251 * 247 * 241 * 239 * 233 * 229;
constexpr double kOzP8hi = (double)kOzP8;
constexpr double kOzP8lo = (double)(long long)(kOzP8 - (unsigned long long)kOzP8hi);
__constant__ int c_oz_m[kOzMods] = {255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199, 197, 193, 191};
constexpr double kOzLog2Half = 123.95306498771728;  // log2((M - 1) / 2), M = product of the moduli

// Offset of element (r, k) of a residue plane in the tiled layout.
__device__ inline int64_t oz_offset(int64_t r, int64_t k, int64_t Kp) {
  return ((r / kOzRows) * (Kp / kOzK) + k / kOzK) * (kOzRows * kOzK) + (r % kOzRows) * kOzK + (k % kOzK);
}

__global__ __launch_bounds__(kOzThreads) void oz_fill_kernel(int* p, int64_t len, int v) {
  for (int64_t e = (int64_t)blockIdx.x * kOzThreads + threadIdx.x; e < len; e += (int64_t)gridDim.x * kOzThreads)
    p[e] = v;
}

// Source element (r, k) of an operand: src[k + r ld] when KMAJ, else src[r + k ld].
template <int KMAJ>
__device__ inline void oz_load_tile(double (*tile)[kOzTile + 1], const double* S, int64_t ld, int64_t R, int64_t K,
                                    int64_t r0, int64_t k0) {
  const int t = threadIdx.x;
#pragma unroll
  for (int it = 0; it < kOzTile * kOzTile / kOzThreads; ++it) {
    const int idx = t + it * kOzThreads;
    const int rr = KMAJ ? idx / kOzTile : idx % kOzTile;
    const int kk = KMAJ ? idx % kOzTile : idx / kOzTile;
    const int64_t r = r0 + rr, k = k0 + kk;
    tile[rr][kk] = (r < R && k < K) ? (KMAJ ? S[k + r * ld] : S[r + k * ld]) : 0.0;
  }
}

// emax[r] = max over k of ilogb|x(r, k)| (INT_MIN for a zero row, INT_MAX if
// the row holds a non-finite value); emax starts at INT_MIN.
template <int KMAJ>
__global__ __launch_bounds__(kOzThreads) void oz_exp_kernel(const double* __restrict__ src, int64_t ld, int64_t stride,
                                                            int64_t R, int64_t K, int* emax, int64_t estride) {
  __shared__ double tile[kOzTile][kOzTile + 1];
  const int64_t z = blockIdx.z;
  oz_load_tile<KMAJ>(tile, src + z * stride, ld, R, K, (int64_t)blockIdx.y * kOzTile, (int64_t)blockIdx.x * kOzTile);
  __syncthreads();
  // 4 threads per row, 16 k each
  const int rr = threadIdx.x >> 2, kk = (threadIdx.x & 3) * 16;
  int e = INT_MIN;
#pragma unroll
  for (int c = 0; c < 16; ++c) {
    const double x = tile[rr][kk + c];
    if (!isfinite(x)) e = INT_MAX;
    else if (x != 0.0 && e != INT_MAX) e = max(e, ilogb(x));
  }
  e = max(e, __shfl_xor(e, 1, 64));
  e = max(e, __shfl_xor(e, 2, 64));
  const int64_t r = (int64_t)blockIdx.y * kOzTile + rr;
  if ((threadIdx.x & 3) == 0 && r < R && e != INT_MIN) atomicMax(emax + z * estride + r, e);
}

// Symmetric residue of an integer a held exactly in a double (|a| < 2^62)
// modulo m, in FP64: q = floor(a / m) from the rounded reciprocal is off by at
// most a few units, and each fma is exact (its exact result is an integer of
// a few times m); the second reduction brings the remainder into [0, m] (m
// itself only when it is an exact multiple and the product rounded down), and
// one comparison into [-h, h] (h = (m - 1) / 2; m maps to 0).  Full-rate FP64
// operations instead of 32-bit integer divisions by constants.
__device__ inline int oz_residue_f64(double a, double m, double inv_m, int h) {
  double r = fma(-floor(a * inv_m), m, a);
  r = fma(-floor(r * inv_m), m, r);
  const int v = (int)r;
  return v > h ? v - (int)m : v;
}

// Residue planes of an operand: row r scaled by 2^(bits - 1 - emax[r]) and
// rounded to an integer, then reduced modulo each modulus (int8, symmetric),
// in the tiled layout, zero outside R x K (rows padded to Rp, depth to Kp).
// Plane l of batch item z starts at out + (z * kOzMods + l) * plane.
template <int KMAJ>
__global__ __launch_bounds__(kOzThreads) void oz_split_kernel(const double* __restrict__ src, int64_t ld, int64_t stride,
                                                              int64_t R, int64_t K, int64_t Kp, const int* emax,
                                                              int64_t estride, int bits, signed char* __restrict__ out,
                                                              int64_t plane) {
  __shared__ double tile[kOzTile][kOzTile + 1];
  const int64_t z = blockIdx.z;
  const int64_t k0 = (int64_t)blockIdx.x * kOzTile, r0 = (int64_t)blockIdx.y * kOzTile;
  oz_load_tile<KMAJ>(tile, src + z * stride, ld, R, K, r0, k0);
  __syncthreads();
  const int rr = threadIdx.x >> 2, kk = (threadIdx.x & 3) * 16;
  const int64_t r = r0 + rr;
  const int em = r < R ? emax[z * estride + r] : INT_MIN;
  const bool live = em != INT_MIN && em != INT_MAX;
  const int sc = live ? bits - 1 - em : 0;
  // the scaled integers, exactly (x 2^sc is exact and rint rounds it to an
  // integer, which a double holds exactly below 2^62)
  double av[16];
#pragma unroll
  for (int c = 0; c < 16; ++c) av[c] = live ? rint(ldexp(tile[rr][kk + c], sc)) : 0.0;
  const int64_t o = oz_offset(r, k0 + kk, Kp);
  signed char* dst = out + z * kOzMods * plane + o;
#pragma unroll
  for (int l = 0; l < kOzMods; ++l) {
    unsigned w[4] = {0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < 16; ++c) {
      const int v = oz_residue_f64(av[c], (double)kOz.m[l], kOz.inv_m[l], kOz.m[l] >> 1);
      w[c >> 2] |= ((unsigned)v & 0xffu) << (8 * (c & 3));
    }
    *reinterpret_cast<uint4*>(dst + l * plane) = make_uint4(w[0], w[1], w[2], w[3]);
  }
}

typedef int i4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// LDS rows of a slab are 128 bytes (8 chunks of 16 bytes), the chunk index
// XOR-swizzled by bits 1..3 of the row: 16 rows x one chunk (a fragment read)
// and 2 rows x 8 chunks (a tile store) each cover the 64 banks once.
__device__ inline int oz_chunk(int row, int c) { return c ^ ((row >> 1) & 7); }

// One int8 GEMM per (batch item, modulus) = blockIdx.z: the residue planes
// of op(A) (rows padded to 256) and op(B) (columns padded to 256), 256 x 256
// workgroup tile of 8 waves (each 128 x 64 = 8 x 4 blocks of 16 x 16), K
// slabs of 128 (two MFMA steps) double-buffered in LDS with a register stage,
// the LDS stores of the staged slab and the loads of the next one spread among
// the MFMAs.  The int32 sums are reduced to [0, m) and stored as bytes
// (column-major M x N planes, leading dimension M).
__global__ __launch_bounds__(kOzGemmThreads) void oz_gemm_kernel(GemmDesc d, const signed char* __restrict__ Ap,
                                                                 const signed char* __restrict__ Bp, int64_t planeA,
                                                                 int64_t planeB, unsigned char* __restrict__ Rp,
                                                                 int64_t planeR, int64_t Kp) {
  __shared__ __attribute__((aligned(16))) signed char LA[2][kOzM][kOzK];  // [stage][row][k, swizzled chunks]
  __shared__ __attribute__((aligned(16))) signed char LB[2][kOzN][kOzK];

  const TileCoord tc = tile_of<kOzM, kOzN>(d.M, d.N, d.lowerC, banded(d), heavy_right(d));
  if (tc.skip) return;
  const int64_t i0 = tc.i0, j0 = tc.j0;
  const int64_t z = blockIdx.z;  // batch item * kOzMods + modulus
  const int mod = c_oz_m[z % kOzMods];
  const signed char* __restrict__ A = Ap + z * planeA;
  const signed char* __restrict__ B = Bp + z * planeB;

  int64_t kb, ke;
  k_range<kOzM, kOzN>(d, i0, j0, kb, ke);
  kb = (kb / kOzK) * kOzK;
  ke = min((ke + kOzK - 1) / kOzK * kOzK, Kp);

  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int wm = wave >> 2, wn = wave & 3;

  i4 acc[8][4];
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = i4{0, 0, 0, 0};

  // A and B slabs: 2 layout tiles each of 1024 chunks (16 B); thread t takes
  // chunks t and t + 512 of each tile.  Chunk e of a tile is row e / 8, k (e % 8) * 16.
  u32x4 a0, a1, a2, a3, b0, b1, b2, b3;
  const int64_t tiles_k = Kp / kOzK;
  const int64_t tile_elems = tiles_k * (kOzRows * kOzK);  // one 128-row block over all of K
  const int64_t oa0 = (i0 / kOzRows) * tile_elems + (int64_t)t * 16;
  const int64_t ob0 = (j0 / kOzRows) * tile_elems + (int64_t)t * 16;
  constexpr int64_t kHalf = 512 * 16;  // chunks t + 512
  auto load_tiles = [&](int64_t k0) {
    const int64_t ko = (k0 / kOzK) * (kOzRows * kOzK);
    a0 = *reinterpret_cast<const u32x4*>(A + oa0 + ko);
    a1 = *reinterpret_cast<const u32x4*>(A + oa0 + ko + kHalf);
    a2 = *reinterpret_cast<const u32x4*>(A + oa0 + tile_elems + ko);
    a3 = *reinterpret_cast<const u32x4*>(A + oa0 + tile_elems + ko + kHalf);
    b0 = *reinterpret_cast<const u32x4*>(B + ob0 + ko);
    b1 = *reinterpret_cast<const u32x4*>(B + ob0 + ko + kHalf);
    b2 = *reinterpret_cast<const u32x4*>(B + ob0 + tile_elems + ko);
    b3 = *reinterpret_cast<const u32x4*>(B + ob0 + tile_elems + ko + kHalf);
  };
  // rows t/8 and 64 + t/8 of each tile (bits 1..3 of the row are equal for
  // rows r, 64 + r, 128 + r and 192 + r)
  const int lrow = t >> 3, lk = oz_chunk(lrow, t & 7) * 16;
  auto store_tiles = [&](int b) {
    *reinterpret_cast<u32x4*>(&LA[b][lrow][lk]) = a0;
    *reinterpret_cast<u32x4*>(&LA[b][64 + lrow][lk]) = a1;
    *reinterpret_cast<u32x4*>(&LA[b][kOzRows + lrow][lk]) = a2;
    *reinterpret_cast<u32x4*>(&LA[b][kOzRows + 64 + lrow][lk]) = a3;
    *reinterpret_cast<u32x4*>(&LB[b][lrow][lk]) = b0;
    *reinterpret_cast<u32x4*>(&LB[b][64 + lrow][lk]) = b1;
    *reinterpret_cast<u32x4*>(&LB[b][kOzRows + lrow][lk]) = b2;
    *reinterpret_cast<u32x4*>(&LB[b][kOzRows + 64 + lrow][lk]) = b3;
  };
  int64_t knext = 0;
  auto slab = [&](int b, bool st, bool ld) {
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      // row bits 1..3 come from lane & 15 alone (the block offsets are multiples of 16)
      const int kc = oz_chunk(lane & 15, ks * 4 + (lane >> 4)) * 16;
      i4 bf[4], af[8];
#pragma unroll
      for (int n = 0; n < 4; ++n)
        bf[n] = *reinterpret_cast<const i4*>(&LB[b][wn * 64 + n * 16 + (lane & 15)][kc]);
#pragma unroll
      for (int m = 0; m < 8; ++m)
        af[m] = *reinterpret_cast<const i4*>(&LA[b][wm * 128 + m * 16 + (lane & 15)][kc]);
      // B fragment as the row operand: the accumulators hold C^T blocks
#pragma unroll
      for (int m = 0; m < 8; ++m)
#pragma unroll
        for (int n = 0; n < 4; ++n) acc[m][n] = __builtin_amdgcn_mfma_i32_16x16x64_i8(bf[n], af[m], acc[m][n], 0, 0, 0);
      if (ks == 0) {
        if (st) store_tiles(b ^ 1);
        if (ld) load_tiles(knext);
      }
    }
    // instruction order: 12 fragment reads; 4 x (8 MFMAs, 2 LDS stores, 2
    // global loads); 12 fragment reads; 32 MFMAs
    __builtin_amdgcn_sched_group_barrier(0x100, 12, 0);  // LDS reads
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      __builtin_amdgcn_sched_group_barrier(0x008, 8, 0);          // MFMA
      if (st) __builtin_amdgcn_sched_group_barrier(0x200, 2, 0);  // LDS write
      if (ld) __builtin_amdgcn_sched_group_barrier(0x020, 2, 0);  // global load
    }
    __builtin_amdgcn_sched_group_barrier(0x100, 12, 0);
    __builtin_amdgcn_sched_group_barrier(0x008, 32, 0);
  };

  if (kb < ke) {
    load_tiles(kb);
    store_tiles(0);
    if (kb + kOzK < ke) load_tiles(kb + kOzK);
    lds_barrier();
    int buf = 0;
    int64_t k0 = kb;
    for (; k0 + 2 * kOzK < ke; k0 += kOzK) {
      knext = k0 + 2 * kOzK;
      slab(buf, true, true);
      lds_barrier();
      buf ^= 1;
    }
    if (k0 + kOzK < ke) {
      slab(buf, true, false);
      lds_barrier();
      buf ^= 1;
    }
    slab(buf, false, false);
  }

  // lane layout of the outputs (C^T blocks): row lane%16, column 4*(lane/16) + v
  unsigned char* __restrict__ R = Rp + z * planeR;
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int64_t row = i0 + wm * 128 + m * 16 + (lane & 15);
    if (row >= d.M) continue;
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      const int64_t colb = j0 + wn * 64 + n * 16 + (lane >> 4) * 4;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        if (colb + v < d.N) {
          int r = acc[m][n][v] % mod;
          if (r < 0) r += mod;
          R[row + (colb + v) * d.M] = (unsigned char)r;
        }
      }
    }
  }
}

// C = alpha_z X 2^-(sA_i + sB_j) + beta C from the residue planes.  X is
// rebuilt from its mixed-radix digits v_l (symmetric, so X lands in
// (-M/2, M/2)): v_l = (r_l - (v_0 + m_0 v_1 + ... ) mod m_l) / (m_0 ... m_{l-1})
// mod m_l, the bracket as one integer multiply-add sum over the table of place
// values mod m_l; then X = L + H m_0...m_7 with the two 8-digit halves L and
// H evaluated exactly in 64-bit integers (|L|, |H| < 2^63), and X as a double
// from one FMA chain (relative error a few units in the last place: when H !=
// 0, |X| > m_0...m_7 / 2 > |L|).  lowerC: only i >= j is written.
__global__ __launch_bounds__(kOzThreads) void oz_crt_kernel(GemmDesc d, const unsigned char* __restrict__ Rp,
                                                            int64_t planeR, const int* __restrict__ emaxA,
                                                            const int* __restrict__ emaxB, int64_t estrideA,
                                                            int64_t estrideB, int bits) {
  const int64_t z = blockIdx.y;
  const int64_t mn = d.M * d.N;
  const double alpha = d.alphaVec ? d.alpha * d.alphaVec[z] : d.alpha;
  double* C = reinterpret_cast<double*>(d.C) + z * d.strideC;
  const unsigned char* R = Rp + z * kOzMods * planeR;
  for (int64_t e = (int64_t)blockIdx.x * kOzThreads + threadIdx.x; e < mn; e += (int64_t)gridDim.x * kOzThreads) {
    const int64_t i = e % d.M, j = e / d.M;
    if (d.lowerC && i < j) continue;
    const int ea = emaxA[z * estrideA + i], eb = emaxB[z * estrideB + j];
    double x;
    if (ea == INT_MAX || eb == INT_MAX) {
      x = __builtin_nan("");
    } else if (ea == INT_MIN || eb == INT_MIN) {
      x = 0.0;
    } else {
      int v[kOzMods];
#pragma unroll
      for (int l = 0; l < kOzMods; ++l) {
        int acc = 0;  // |acc| <= 15 * 127 * 254 < 2^19
#pragma unroll
        for (int q = 0; q < l; ++q) acc += v[q] * kOz.pre[l][q];
        int t = ((int)R[l * planeR + e] - acc) % kOz.m[l];
        if (t < 0) t += kOz.m[l];
        t = t * kOz.qinv[l] % kOz.m[l];
        v[l] = t > (kOz.m[l] >> 1) ? t - kOz.m[l] : t;
      }
      long long L = v[7], H = v[15];
#pragma unroll
      for (int l = 6; l >= 0; --l) L = L * kOz.m[l] + v[l];
#pragma unroll
      for (int l = 14; l >= 8; --l) H = H * kOz.m[l] + v[l];
      const double hd = (double)H;
      x = fma(hd, kOzP8hi, fma(hd, kOzP8lo, (double)L));
      x = ldexp(x, -(2 * (bits - 1) - ea - eb));
    }
    double* p = C + i + j * d.ldc;
    const double val = alpha * x;
    *p = (d.beta == 0.0) ? val : val + d.beta * (*p);
  }
}

void* g_oz_ws = nullptr;
size_t g_oz_cap = 0;

}  // namespace

bool gemm_ozaki_supported(const GemmDesc& d) {
  // one operand per batch item (no maps), no split-K, K small enough for the
  // moduli to cover the products at 53 bits or more
  return !d.mapA && !d.mapB && !d.mapC && d.splitk <= 1 && d.K >= 1 && d.K <= 131072;
}

// The op(A) whose residues and exponents the workspace holds (set after a
// conversion of a single-item product, cleared by any other conversion of A
// or a new workspace).
struct OzLastA {
  bool valid = false;
  const void* A = nullptr;
  void* ws = nullptr;
  int64_t lda = 0, M = 0, N = 0, K = 0;
  int transA = 0, bits = 0;
};
OzLastA g_oz_last;

int gemm_ozaki(const GemmDesc& d, hipStream_t s, bool capture) {
  if (!gemm_ozaki_supported(d)) {
    set_error("gemm (int8 moduli): unsupported descriptor");
    return kErrArg;
  }
  const int bits = std::min(60, (int)std::floor((kOzLog2Half - std::log2((double)d.K)) / 2.0));
  const int64_t Mp = cdiv(d.M, kOzM) * kOzM, Np = cdiv(d.N, kOzN) * kOzN, Kp = cdiv(d.K, kOzK) * kOzK;
  const int64_t planeA = Mp * Kp, planeB = Np * Kp, planeR = d.M * d.N;
  // one batch item at a time: 16 (Mp + Np) Kp + 16 M N bytes of residues
  const size_t bytesA = (size_t)kOzMods * planeA, bytesB = (size_t)kOzMods * planeB;
  const size_t bytesR = (size_t)kOzMods * planeR;
  const size_t offE = (bytesA + bytesB + bytesR + 255) / 256 * 256;
  const size_t bytes = offE + sizeof(int) * (size_t)(Mp + Np);
  if (bytes > g_oz_cap) {
    if (capture) {
      set_error("gemm (int8 moduli): the residue workspace cannot grow during graph capture");
      return kErrArg;
    }
    if (g_oz_ws) {
      ZOLO_HIP(hipDeviceSynchronize());
      ZOLO_HIP(hipFree(g_oz_ws));
    }
    g_oz_ws = nullptr;
    g_oz_cap = 0;
    g_oz_last.valid = false;  // a new allocation may reuse the old address
    ZOLO_HIP(hipMalloc(&g_oz_ws, bytes));
    g_oz_cap = bytes;
  }
  signed char* RA = reinterpret_cast<signed char*>(g_oz_ws);
  unsigned char* RC = reinterpret_cast<unsigned char*>(RA + bytesA + bytesB);
  int* eA = reinterpret_cast<int*>(reinterpret_cast<char*>(g_oz_ws) + offE);
  // op(A) = op(B)^T of one matrix (A^T A or A A^T, e.g. the polish's U^T U):
  // the rows of op(A) and the columns of op(B) are the same vectors with the
  // same scaling, so the residues of op(A) serve as those of op(B)
  const bool same = d.A == d.B && d.lda == d.ldb && d.strideA == d.strideB && d.M == d.N && d.transA != d.transB;
  signed char* RB = same ? RA : RA + bytesA;
  int* eB = same ? eA : eA + Mp;
  const dim3 ga((unsigned)(Kp / kOzTile), (unsigned)(Mp / kOzTile), 1);
  const dim3 gb((unsigned)(Kp / kOzTile), (unsigned)(Np / kOzTile), 1);
  const int64_t tiles = cdiv(d.M, kOzM) * cdiv(d.N, kOzN);
  const unsigned gx = (unsigned)std::min<int64_t>(8192, std::max<int64_t>(1, cdiv(planeR, kOzThreads)));
  for (int b = 0; b < d.batch; ++b) {
    const double* A = reinterpret_cast<const double*>(d.A) + (int64_t)b * d.strideA;
    const double* B = reinterpret_cast<const double*>(d.B) + (int64_t)b * d.strideB;
    GemmDesc db = d;
    db.batch = 1;
    db.C = reinterpret_cast<double*>(d.C) + (int64_t)b * d.strideC;
    db.alphaVec = d.alphaVec ? d.alphaVec + b : nullptr;
    // an operand shared by the batch (stride 0) is converted once
    // (the layout of the workspace depends on M, N and K only, so a match of
    // those, the operand and the bits finds the residues where they were left)
    const bool keepA = d.sameA && d.batch == 1 && g_oz_last.valid && g_oz_last.A == d.A && g_oz_last.ws == g_oz_ws &&
                       g_oz_last.lda == d.lda && g_oz_last.M == d.M && g_oz_last.N == d.N && g_oz_last.K == d.K &&
                       g_oz_last.transA == d.transA && g_oz_last.bits == bits;
    const bool convA = !keepA && (b == 0 || d.strideA != 0), convB = !same && (b == 0 || d.strideB != 0);
    if (convA) {
      g_oz_last.valid = d.batch == 1;
      g_oz_last.A = d.A;
      g_oz_last.ws = g_oz_ws;
      g_oz_last.lda = d.lda;
      g_oz_last.M = d.M;
      g_oz_last.N = d.N;
      g_oz_last.K = d.K;
      g_oz_last.transA = d.transA;
      g_oz_last.bits = bits;
    }
    if (convA) {
      hipLaunchKernelGGL(oz_fill_kernel, dim3((unsigned)cdiv(Mp, kOzThreads)), dim3(kOzThreads), 0, s, eA, Mp, INT_MIN);
      // op(A) row i, depth k: A[k + i lda] when transposed (K-major), else A[i + k lda]
      if (d.transA) {
        hipLaunchKernelGGL((oz_exp_kernel<1>), ga, dim3(kOzThreads), 0, s, A, d.lda, (int64_t)0, d.M, d.K, eA, Mp);
        hipLaunchKernelGGL((oz_split_kernel<1>), ga, dim3(kOzThreads), 0, s, A, d.lda, (int64_t)0, d.M, d.K, Kp, eA,
                           Mp, bits, RA, planeA);
      } else {
        hipLaunchKernelGGL((oz_exp_kernel<0>), ga, dim3(kOzThreads), 0, s, A, d.lda, (int64_t)0, d.M, d.K, eA, Mp);
        hipLaunchKernelGGL((oz_split_kernel<0>), ga, dim3(kOzThreads), 0, s, A, d.lda, (int64_t)0, d.M, d.K, Kp, eA,
                           Mp, bits, RA, planeA);
      }
    }
    if (convB) {
      hipLaunchKernelGGL(oz_fill_kernel, dim3((unsigned)cdiv(Np, kOzThreads)), dim3(kOzThreads), 0, s, eB, Np, INT_MIN);
      // op(B) column j, depth k: B[k + j ldb] when not transposed (K-major), else B[j + k ldb]
      if (d.transB) {
        hipLaunchKernelGGL((oz_exp_kernel<0>), gb, dim3(kOzThreads), 0, s, B, d.ldb, (int64_t)0, d.N, d.K, eB, Np);
        hipLaunchKernelGGL((oz_split_kernel<0>), gb, dim3(kOzThreads), 0, s, B, d.ldb, (int64_t)0, d.N, d.K, Kp, eB,
                           Np, bits, RB, planeB);
      } else {
        hipLaunchKernelGGL((oz_exp_kernel<1>), gb, dim3(kOzThreads), 0, s, B, d.ldb, (int64_t)0, d.N, d.K, eB, Np);
        hipLaunchKernelGGL((oz_split_kernel<1>), gb, dim3(kOzThreads), 0, s, B, d.ldb, (int64_t)0, d.N, d.K, Kp, eB,
                           Np, bits, RB, planeB);
      }
    }
    ZOLO_HIP(hipGetLastError());
    hipLaunchKernelGGL(oz_gemm_kernel, dim3((unsigned)tiles, 1, kOzMods), dim3(kOzGemmThreads), 0, s, db, RA, RB,
                       planeA, planeB, RC, planeR, Kp);
    ZOLO_HIP(hipGetLastError());
    hipLaunchKernelGGL(oz_crt_kernel, dim3(gx, 1), dim3(kOzThreads), 0, s, db, RC, planeR, eA, eB, (int64_t)0,
                       (int64_t)0, bits);
    ZOLO_HIP(hipGetLastError());
  }
  return kOk;
}

}  // namespace zolo
