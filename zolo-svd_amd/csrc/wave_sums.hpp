// Column sums over the 64 lanes of a wave: every lane holds one term of each
// of C columns (a register array) and lane l receives the sum of column l
// (l & 31 for 32 columns).  A butterfly in which each stage exchanges half of
// the remaining columns with the partner lane: C - 1 exchanges instead of
// C x 6.  All indices are compile-time, so the arrays stay in registers.
#pragma once
#include <hip/hip_runtime.h>

namespace zolo {

// The partner lane of a butterfly stage of width H <= 8, read with a DPP
// move (no LDS round trip): row_mirror (lane i of a 16-lane row reads lane
// 15 - i) for H = 8, row_half_mirror (7 - i within 8 lanes) for H = 4, the
// quad permutations [2,3,0,1] and [1,0,3,2] for H = 2 and 1.  Each pairs a
// lane with one whose bit H differs and whose higher bits agree, which is all
// the butterfly needs: after the four stages lane l holds the sum over its
// 16-lane row of column l & 15, as with lane ^ H partners.
template <int H>
constexpr int kDppPair = H == 8 ? 0x140 : H == 4 ? 0x141 : H == 2 ? 0x4E : 0xB1;
template <int H>
__device__ inline unsigned dpp_pair(unsigned x) {
  return (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, kDppPair<H>, 0xF, 0xF, true);
}
template <int H>
__device__ inline double lane_pair(double x) {
  if constexpr (H <= 8) {
    return __hiloint2double((int)dpp_pair<H>((unsigned)__double2hiint(x)),
                            (int)dpp_pair<H>((unsigned)__double2loint(x)));
  } else {
    return __shfl_xor(x, H, 64);
  }
}
template <int H>
__device__ inline float lane_pair(float x) {
  if constexpr (H <= 8) return __uint_as_float(dpp_pair<H>(__float_as_uint(x)));
  else return __shfl_xor(x, H, 64);
}

// One shuffle stage: lanes with bit H set keep columns j + H, the others j.
template <int H, typename V>
__device__ inline void butterfly_step(V* a, int lane) {
  const bool up = (lane & H) != 0;
#pragma unroll
  for (int j = 0; j < H; ++j) {
    const V keep = up ? a[j + H] : a[j];
    const V send = up ? a[j] : a[j + H];
    a[j] = keep + lane_pair<H>(send);
  }
}
template <typename V>
__device__ inline V wave_colsum64(V* a, int lane) {
  butterfly_step<32>(a, lane);
  butterfly_step<16>(a, lane);
  butterfly_step<8>(a, lane);
  butterfly_step<4>(a, lane);
  butterfly_step<2>(a, lane);
  butterfly_step<1>(a, lane);
  return a[0];
}

// One butterfly stage over lane halves (W = 32) or 16-lane rows (W = 16) with
// the lane-swap instructions: permlane32_swap exchanges lanes 32..63 of X with
// lanes 0..31 of Y (permlane16_swap the odd rows of X with the even rows of
// Y), so X + Y afterwards holds column "X" in the lower and column "Y" in the
// upper lanes, summed over each pair of lanes W apart.
template <int W>
__device__ inline void lane_swap(unsigned& x, unsigned& y) {
  if constexpr (W == 32) {
    const auto r = __builtin_amdgcn_permlane32_swap(x, y, false, false);
    x = r[0];
    y = r[1];
  } else {
    const auto r = __builtin_amdgcn_permlane16_swap(x, y, false, false);
    x = r[0];
    y = r[1];
  }
}
template <int W>
__device__ inline double swap_add(double X, double Y) {
  unsigned xl = __double2loint(X), xh = __double2hiint(X);
  unsigned yl = __double2loint(Y), yh = __double2hiint(Y);
  lane_swap<W>(xl, yl);
  lane_swap<W>(xh, yh);
  return __hiloint2double(xh, xl) + __hiloint2double(yh, yl);
}
template <int W>
__device__ inline float swap_add(float X, float Y) {
  unsigned x = __float_as_uint(X), y = __float_as_uint(Y);
  lane_swap<W>(x, y);
  return __uint_as_float(x) + __uint_as_float(y);
}

// Column sums over the 64 lanes of the wave for 32 columns, a[c] (c < 32)
// holding each lane's term of column c (a is overwritten): the first stage
// with the 16-lane swap, then four shuffle stages; lane l returns the sum of
// column l & 31 (the last stage adds the two lane halves with the 32-lane swap).
template <typename V>
__device__ inline V wave_colsum32_swap(V* a, int lane) {
#pragma unroll
  for (int j = 0; j < 16; ++j) a[j] = swap_add<16>(a[j], a[j + 16]);
  butterfly_step<8>(a, lane);
  butterfly_step<4>(a, lane);
  butterfly_step<2>(a, lane);
  butterfly_step<1>(a, lane);
  return swap_add<32>(a[0], a[0]);
}

// Column sums over the 64 lanes for 16 columns, a[c] (c < 16) holding each
// lane's term of column c (a is overwritten): four shuffle stages inside each
// 16-lane row, then the four rows added with the 16- and 32-lane swaps; lane
// l returns the sum of column l & 15.
template <typename V>
__device__ inline V wave_colsum16_swap(V* a, int lane) {
  butterfly_step<8>(a, lane);
  butterfly_step<4>(a, lane);
  butterfly_step<2>(a, lane);
  butterfly_step<1>(a, lane);
  const V r = swap_add<16>(a[0], a[0]);
  return swap_add<32>(r, r);
}

// Column sums of x * v[c] over the 64 lanes for all 64 columns: lane l
// returns the sum of column l.  The first two stages use the lane swaps, the
// last four shuffles; the products are formed in the first stage.
template <typename V>
__device__ inline V wave_colsum64_swap(const V* v, V x, int lane) {
  // the first two stages fused per j, so only 16 partial columns are live
  V a[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const V lo = swap_add<32>(x * v[j], x * v[j + 32]);
    const V hi = swap_add<32>(x * v[j + 16], x * v[j + 48]);
    a[j] = swap_add<16>(lo, hi);
  }
  butterfly_step<8>(a, lane);
  butterfly_step<4>(a, lane);
  butterfly_step<2>(a, lane);
  butterfly_step<1>(a, lane);
  return a[0];
}

}  // namespace zolo
