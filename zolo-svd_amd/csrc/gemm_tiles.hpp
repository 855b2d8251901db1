// Tile mapping, K-range trimming and LDS helpers shared by the MFMA GEMM
// kernels (gemm.hip, ozaki.hip).
#pragma once
#include "common.hpp"

namespace zolo {

// Workgroup barrier that first waits for this wave's LDS operations only:
// global loads in flight stay in flight across it.  The memory clobber keeps
// the compiler from moving LDS accesses across the barrier.
__device__ inline void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

struct TileCoord { int64_t i0, j0; bool skip; };

// Map the linear workgroup id to a (row, col) tile, XCD-aware and grouped.
// XCD x (= pid % 8, the round-robin dispatch) owns a contiguous eighth of the
// logical tiles, which walk groups of 8 tile-rows column by column so that
// neighbouring workgroups share A and B panels in the XCD's L2.  When the work
// per tile varies along the rows (triangular outputs, banded operands),
// "balance" orders the full row groups 0, last, 1, last-1, ... (with groups
// small enough that every XCD gets at least two) so that every XCD's eighth
// pairs heavy and light groups.  When the work grows with the tile column
// (rev_cols: op(B) zero below a diagonal, K = [0, j0 + BN + bHi)), each group
// walks its columns last to first, so the heaviest tiles are dispatched first
// and the lightest make up the tail.
template <int BM, int BN>
__device__ inline TileCoord tile_of(int64_t M, int64_t N, int lowerC, bool balance, bool rev_cols = false) {
  const int64_t tm = (M + BM - 1) / BM, tn = (N + BN - 1) / BN;
  const int64_t nt = tm * tn;
  int64_t pid = blockIdx.x;
  const int64_t full = (nt / 8) * 8;
  if (pid < full) {  // XCD x (= pid % 8) owns logical tiles [x*nt/8, (x+1)*nt/8)
    pid = (pid % 8) * (full / 8) + pid / 8;
  }
  // tile-rows per group; balancing wants at least two full groups per XCD
  const int64_t G = balance ? max((int64_t)1, min((int64_t)8, tm / 16)) : 8;
  int64_t group = pid / (G * tn);
  const int64_t r = pid - group * G * tn;
  const int64_t nfull = tm / G;  // groups of exactly G tile-rows (a partial one stays last)
  if (balance && group < nfull) group = (group & 1) ? nfull - 1 - (group >> 1) : (group >> 1);
  const int64_t first = group * G;
  const int64_t gsz = (tm - first) < G ? (tm - first) : G;
  TileCoord t;
  t.i0 = (first + r % gsz) * BM;
  t.j0 = (rev_cols ? tn - 1 - r / gsz : r / gsz) * BN;
  t.skip = lowerC && (t.j0 > t.i0 + BM - 1);
  return t;
}

// Work per tile depends on its position: triangular output or trimmed K range.
__device__ inline bool banded(const GemmDesc& d) {
  return d.lowerC || d.aLo >= 0 || d.aHi >= 0 || d.bLo >= 0 || d.bHi >= 0;
}
// The K range grows with the tile column and nothing else trims it.
__device__ inline bool heavy_right(const GemmDesc& d) {
  return d.bHi >= 0 && d.bLo < 0 && d.aLo < 0 && d.aHi < 0 && !d.lowerC;
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

// LDS rows of a slab are 64 bytes (4 chunks of 16 bytes) with the chunk index
// XOR-swizzled by bits 1..2 of the row: the 16-byte fragment reads (16 rows x
// 4 chunks per wave) and the 16-byte tile writes (16 rows x 4 chunks) then
// touch 64 distinct banks per 256-byte LDS cycle.  A padded 80-byte pitch
// conflicts two ways on both (counter probe: 8 instead of 4 LDS cycles per
// wave-wide 16-byte read, 16 instead of 8 per write).
__device__ inline int lds_chunk(int row, int c) { return c ^ ((row >> 1) & 3); }

}  // namespace zolo
