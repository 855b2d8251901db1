// Structured Householder QR of the stacked matrix [X; s I] (2n x n), batched
// over the Zolotarev shifts (blockIdx.y), following mpdgeqrf.f / mpdorgqr.f:
// the reflector of column j only has support on rows j .. n+j, because the
// identity block below X fills in one row per eliminated column.
//
// A panel at column p of width jb works on two row ranges, [p, r1_end) of the
// top block and [n, r2_end) of the bottom block; column j of the panel only
// involves the bottom rows up to n + j.  [X; s I]: r1_end = n, r2_end =
// n + p + jb (one contiguous range).  A dense n x n matrix: r1_end = n, no
// bottom rows.  [R; s I] with R upper triangular: r1_end = p + jb (the top
// rows below the panel are zero and stay zero), r2_end = n + p + jb.
//
// Panel factorisation, three forms: the whole panel in one workgroup for
// stacks of at most kQrPanelWgRows rows (qr_panel_wg_kernel); the whole panel
// in one launch whose workgroups exchange one column sum per column through
// per-workgroup flags (qr_panel_coop_kernel); and, where neither applies (graph capture of a
// multi-workgroup panel, grids too large to be resident), one launch per
// column (qr_panel_kernel): launch j applies the reflector of column j-1 to
// the rest of the panel and, in the same pass over the rows each thread owns,
// reduces the Gram row g_c = sum_i S[i][j] S[i][c] that the reflector of
// column j needs (its norm is g_j).  Partial sums go to per-workgroup slots
// and are summed in a fixed order, so the result does not depend on
// scheduling.  The trailing update and the explicit Q use the compact WY form
// (I - V T V^T) through the MFMA GEMM.
#include <algorithm>
#include <type_traits>

#include "kernels.hpp"
#include "wave_sums.hpp"

namespace zolo {
namespace {

constexpr int kT = 256;
constexpr int NBP = kQrNb;

// part layout per shift: for local column jj, a record of (nwg*jb + jb) doubles:
// [0, nwg*jb): per-workgroup Gram partials, [nwg*jb, nwg*jb+jb): pivot row.
// Each thread owns one row of the panel and keeps its jb values in registers
// for the whole launch: one batch of loads, the rank-1 update and the Gram
// products in registers, one batch of stores.
template <typename T>
__global__ __launch_bounds__(kT) void qr_panel_kernel(T* S, int64_t lds, int64_t strideS, int64_t n, int64_t p, int jb,
                                                      int64_t r1_end, int64_t r2_end, int64_t j, int apply_prev,
                                                      int reduce_cur, double* part, int nwg, T* tau,
                                                      int64_t strideTau) {
  __shared__ double gs[NBP], rr[NBP], wv[NBP];
  __shared__ double sc[3];  // beta, tau, scal
  __shared__ double red[4][NBP];
  const int q = blockIdx.y;
  T* Sq = S + q * strideS;
  const int64_t rec = (int64_t)nwg * jb + jb;
  double* base = part + (int64_t)q * jb * rec;
  const int t = threadIdx.x, w = blockIdx.x;
  // the row this thread owns: local index l over [p, r1_end) then [n, r2_end)
  const int64_t l = (int64_t)w * kT + t, len1 = r1_end - p;
  const int64_t i = l < len1 ? p + l : n + (l - len1);
  const int64_t jp = j - 1;
  const int jjp = (int)(jp - p), jj = (int)(j - p);
  // bottom rows beyond n + j take no part in either reflector of this launch
  const bool live = l < len1 + (r2_end - n) && (i < r1_end || i <= n + j) && i >= (apply_prev ? jp : j);

  // only the columns this launch updates or reduces: [jjp, jb) or [jj, jb)
  const int cmin = apply_prev ? jjp : jj;
  double v[NBP];
#pragma unroll
  for (int c = 0; c < NBP; ++c) v[c] = (live && c >= cmin && c < jb) ? (double)Sq[i + (p + c) * lds] : 0.0;

  if (apply_prev) {
    const double* G = base + (int64_t)jjp * rec;
    {  // Gram partials of the nwg workgroups: 4 groups of 64 threads, 4 loads in flight each
      const int c = t & 63, g = t >> 6;
      double s = 0.0;
      if (c < jb && c >= jjp) {
        int k = g;
        for (; k + 12 < nwg; k += 16)
          s += (G[(int64_t)k * jb + c] + G[(int64_t)(k + 4) * jb + c]) +
               (G[(int64_t)(k + 8) * jb + c] + G[(int64_t)(k + 12) * jb + c]);
        for (; k < nwg; k += 4) s += G[(int64_t)k * jb + c];
      }
      red[g][c] = s;
      if (t < jb) rr[t] = G[(int64_t)nwg * jb + t];
    }
    __syncthreads();
    if (t < jb) gs[t] = (red[0][t] + red[1][t]) + (red[2][t] + red[3][t]);
    __syncthreads();
    if (t == 0) {
      const double alpha = rr[jjp], xn2 = gs[jjp];
      double beta = alpha, tv = 0.0, scal = 0.0;
      if (xn2 > 0.0) {
        beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
        tv = (beta - alpha) / beta;
        scal = 1.0 / (alpha - beta);
      }
      sc[0] = beta; sc[1] = tv; sc[2] = scal;
      if (w == 0) tau[q * strideTau + jp] = (T)tv;
    }
    __syncthreads();
    if (t < jb) wv[t] = (t > jjp) ? sc[1] * (rr[t] + sc[2] * gs[t]) : 0.0;
    __syncthreads();
    if (live && i <= n + jp) {
      // v[jjp] without a dynamic register index (which would spill v to scratch)
      double vj = 0.0;
#pragma unroll
      for (int c = 0; c < NBP; ++c)
        if (c == jjp) vj = v[c];
      const double vi = (i == jp) ? 1.0 : vj * sc[2];
      const double newj = (i == jp) ? sc[0] : vi;
#pragma unroll
      for (int c = 0; c < NBP; ++c) {
        if (c == jjp) v[c] = newj;
        if (c > jjp && c < jb) v[c] -= vi * wv[c];
      }
#pragma unroll
      for (int c = 0; c < NBP; ++c)
        if (c >= jjp && c < jb) Sq[i + (p + c) * lds] = (T)v[c];
    }
  }

  if (reduce_cur) {
    double* G = base + (int64_t)jj * rec;
    if (live && i == j) {
#pragma unroll
      for (int c = 0; c < NBP; ++c)
        if (c >= jj && c < jb) G[(int64_t)nwg * jb + c] = v[c];
    }
    double x = 0.0;
#pragma unroll
    for (int c = 0; c < NBP; ++c)
      if (c == jj) x = v[c];
    if (!(live && i > j)) x = 0.0;
    const int lane = t & 63, wave = t >> 6;
#pragma unroll
    for (int c = 0; c < NBP; ++c) v[c] = (c >= jj && c < jb) ? x * v[c] : 0.0;
    red[wave][lane] = wave_colsum64(v, lane);
    __syncthreads();
    if (t < jb) G[(int64_t)w * jb + t] = red[0][t] + red[1][t] + red[2][t] + red[3][t];
  }
}

// Whole panel (jb <= 64 columns) in one launch: one thread per panel row (rows
// [p, r1_end) then [n, r2_end)) keeps the row's panel values in registers,
// shifted so that the current column is always v[0] (no register is selected
// by the loop counter), and stores each column when it is finished.  The
// columns run in three phases by the number of columns still in the panel
// (more than 32, 17 to 32, at most 16): the column sums, the update and the
// shift then touch 64, 32 or 16 registers.
//
// Per column: the wave column sums give the Gram row g_c = sum_{i>j} S[i][j]
// S[i][j+c] (lane 0: the squared norm below the diagonal); after a workgroup
// barrier wave 0 adds the wave sums.  With one workgroup per shift it forms
// the reflector (dlarfg) and w_c = tau (r_c + scal g_c) at once; after a
// second barrier every thread applies the reflector to its row.
//
// Workgroups of TT = 256 threads (one wave per SIMD, for launches that leave
// CUs free) or 512 threads (two waves per SIMD, for grids that fill the
// device).  With several workgroups per shift the sums are exchanged through
// device memory, in one of two ways.  All sums run in a fixed order, so the
// results do not depend on scheduling, and no counter is shared by the
// workgroups.  Slots and flags are only accessed by device-scope atomic loads
// and stores, which gfx950 performs at the device's coherence point (stores
// written through, loads not served from an incoherent cache), so neither way
// needs the L2 writeback and invalidate of a release/acquire pair.
//
// Slot polling (TT = 256): every column has a record of its own, filled with
// kSlotEmpty before the launch (a value the stores never write: a sum with
// that bit pattern is stored as the default NaN).  Wave 0 stores the
// workgroup's sums (workgroup 0 also the pivot row j, its row jj) to the
// column's record; every wave loads its share of the slots (wave w:
// workgroups w, w + 4, ...) and reloads the empty ones until none is left,
// then wave 0 adds the shares.  One store-to-load round trip per column.
//
// Shift per XCD (xmap; at most 8 shifts of at most 32 workgroups): workgroup
// 8 w + qi is row block w of shift qi, so under the round-robin dispatch of
// workgroups to the 8 XCDs all of a shift's run on one XCD; each checks that
// from its XCC id at the start.  Then that XCD's L2 is the shift's coherence
// point: the slot stores are workgroup-scope (sc0: into the L2, not written
// through) and the device-scope polls (sc1) see them there, about 300 ns after
// the store instead of 560-840 ns for a device-scope store (measured with
// one workgroup storing a flag and another on the same XCD polling it;
// workgroup-scope loads would be served by the poller's own L1 and never see
// it).  TT = 256 polls the slots, TT = 512 exchanges by flags (both with
// workgroup-scope stores); 204-222 us per 64-column panel of up to 8192 rows
// against 236-254 us with device-scope stores across XCDs, and 243-266 us
// for 8-shift panels over 8192 rows against about 320 us.
//
// Flags (TT = 512): wave 0 stores the sums to the column's slot, waits until
// those stores are complete (vmcnt(0)), raises the workgroup's flag to jj + 1
// and polls the nwg flags of the shift together (lane k reads flag k) until
// all have reached jj + 1; after a barrier every wave loads its share of the
// slots, and after another wave 0 adds the shares.  Two round trips, but the
// pollers read one line per shift instead of every slot, which matters when
// every CU holds a workgroup that polls.  The slots alternate between two
// records: a workgroup writes column jj's slot only after every workgroup has
// raised its flag for column jj - 1, which each does after it has read the
// slots of column jj - 2.
//
// Shifts q0 .. q0 + gridDim.y - 1; flags holds kCoopFlagStride zeroed words
// per shift (flag exchange) or is null.  A poll that outlasts its spin cap
// gives up (the results are then wrong, but the kernel ends), so a launch
// that is not co-resident cannot hang the device.
constexpr int kCoopMaxWg = 64;        // workgroups per shift: one flag per lane of a wave, 16 slots per wave
constexpr int kCoopFlagStride = 64;   // unsigned words between the flags of two shifts (one 256-byte line)
constexpr unsigned kCoopSpinCap = 1u << 24;
// s_waitcnt operand (gfx9 encoding): vmcnt(0), expcnt and lgkmcnt not waited for
constexpr int kWaitVm0 = 0x0F70;

// Slot value that no partial sum is stored as: the launcher fills the slot
// records with it, and a reader polls a slot until it holds something else.
constexpr long long kSlotEmpty = -1LL;  // all bits set: a NaN
constexpr unsigned kSlotSpinCap = 1u << 20;
__device__ inline bool slot_empty(double x) { return __double_as_longlong(x) == kSlotEmpty; }
__device__ inline double slot_value(double x) {
  return slot_empty(x) ? __longlong_as_double(0x7FF8000000000000LL) : x;
}

template <typename T, int TT>
__global__ __launch_bounds__(TT) void qr_panel_coop_kernel(T* S, int64_t lds, int64_t strideS, int64_t n,
                                                               int64_t p, int jb, int64_t r1_end, int64_t r2_end,
                                                               double* part, unsigned* flags, int nwg, T* tau,
                                                               int64_t strideTau, int q0, int xmap, int nq) {
  constexpr int kCoopWaves = TT / 64;
  // exchange by polling the slots (256 threads) or by flags (512 threads)
  constexpr bool XM = TT <= 256;
  __shared__ double red[kCoopWaves][NBP], red2[kCoopWaves][NBP];
  __shared__ double wv[NBP], prj[NBP];
  __shared__ double sc[2];
  __shared__ int xlocal;
  // xmap: one shift per XCD, workgroup 8 w + qi being row block w of shift
  // qi (the round-robin dispatch puts every eighth workgroup on one XCD);
  // else row block blockIdx.x of shift blockIdx.y
  const int qi = xmap ? (int)(blockIdx.x & 7u) : (int)blockIdx.y;
  const int w = xmap ? (int)(blockIdx.x >> 3) : (int)blockIdx.x;
  if (qi >= nq) return;
  const int q = q0 + qi;
  T* Sq = S + (int64_t)q * strideS;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int64_t l = (int64_t)w * TT + t, len1 = r1_end - p;
  const int64_t i = l < len1 ? p + l : n + (l - len1);
  const bool valid = l < len1 + (r2_end - n);
  const int64_t rec = (int64_t)nwg * NBP + NBP;
  double* base = part + (int64_t)qi * (XM ? NBP : 2) * rec;
  const unsigned* flr = flags ? flags + (int64_t)qi * kCoopFlagStride : nullptr;
  unsigned* flw = flags ? flags + (int64_t)qi * kCoopFlagStride + w : nullptr;
  // xmap with several workgroups: do all of this shift's run on one XCD?  Each
  // publishes 1 + its XCC id (device scope, the words filled with ~0 before
  // the launch) and reads the others'; all decide from the same values.
  // Then that XCD's L2 is the shift's coherence point: slot stores go to it
  // (workgroup scope, sc0) and the device-scope polls (sc1) see them there.
  if (t == 0) {
    int loc = 0;
    if (xmap == 1 && nwg > 1) {
      unsigned x;
      asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(x));
      const int mine = (int)(x & 0xFu) + 1;
      int* xid = reinterpret_cast<int*>(part + (int64_t)nq * (XM ? NBP : 2) * rec) + qi * kCoopMaxWg;
      __hip_atomic_store(xid + w, mine, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      loc = 1;
      unsigned spin = 0;
      for (int k = 0; k < nwg; ++k) {
        int v;
        while ((v = __hip_atomic_load(xid + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) == -1 &&
               ++spin < kSlotSpinCap)
          __builtin_amdgcn_s_sleep(1);
        if (v != mine) loc = 0;
      }
    }
    xlocal = loc;
  }
  __syncthreads();
  const bool l2local = xlocal != 0;
  // v[c]: the row's value in panel column jj + c; zero for jj + c >= jb
  double v[NBP];
#pragma unroll
  for (int c = 0; c < NBP; ++c) v[c] = (valid && c < jb) ? (double)Sq[i + (p + c) * lds] : 0.0;
  T* out = Sq + i + p * lds;

  auto column = [&](auto cols, int jj) {
    constexpr int NV = decltype(cols)::value;
    const int64_t j = p + jj;
    const int rem = jb - jj;  // columns jj .. jb - 1 are v[0 .. rem), rem <= NV
    // bottom rows beyond n + j take no part in the reflector of column j
    const bool live = valid && i >= j && (i < r1_end || i <= n + j);
    // the pivot row j is row jj of the top range: thread jj of workgroup 0 (wave 0)
    if (live && i == j) {
#pragma unroll
      for (int c = 0; c < NV; ++c) prj[c] = v[c];
    }
    const double x = (live && i > j) ? v[0] : 0.0;
    double cs;
    if constexpr (NV == 64) {
      cs = wave_colsum64_swap(v, x, lane);
    } else {
      double a[NV];
#pragma unroll
      for (int c = 0; c < NV; ++c) a[c] = x * v[c];
      if constexpr (NV == 32) cs = wave_colsum32_swap(a, lane);
      else cs = wave_colsum16_swap(a, lane);
    }
    red[wave][lane] = cs;  // lanes >= NV hold copies of lane & (NV - 1), ignored below (rem <= NV)
    __syncthreads();
    double g = 0.0, pr = 0.0;
    if (wave == 0) {
#pragma unroll
      for (int k = 0; k < kCoopWaves; ++k) g += red[k][lane];
      pr = prj[lane];
    }
    if constexpr (XM) {
      if (nwg > 1) {
      // one record per column, filled with kSlotEmpty before the launch:
      // wave 0 stores the workgroup's sums, every wave polls its share of the
      // slots until none is empty
      double* G = base + (int64_t)jj * rec;
      if (wave == 0 && lane < rem) {
        if (l2local) {
          __hip_atomic_store(G + (int64_t)w * NBP + lane, slot_value(g), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_WORKGROUP);
          if (w == 0)
            __hip_atomic_store(G + (int64_t)nwg * NBP + lane, slot_value(pr), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
        } else {
          __hip_atomic_store(G + (int64_t)w * NBP + lane, slot_value(g), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (w == 0)
            __hip_atomic_store(G + (int64_t)nwg * NBP + lane, slot_value(pr), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
      }
      constexpr int M = kCoopMaxWg / kCoopWaves;
      double sv[M];
      const double* Gl = G + lane;
#pragma unroll
      for (int m = 0; m < M; ++m) {
        const int k = wave + m * kCoopWaves;
        sv[m] = (k < nwg && lane < rem) ? __hip_atomic_load(Gl + (int64_t)k * NBP, __ATOMIC_RELAXED,
                                                            __HIP_MEMORY_SCOPE_AGENT)
                                        : 0.0;
      }
      double pv = (wave == 0 && lane < rem)
                      ? __hip_atomic_load(Gl + (int64_t)nwg * NBP, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                      : 0.0;
      for (unsigned spin = 0; spin < kSlotSpinCap; ++spin) {
        bool miss = slot_empty(pv);
#pragma unroll
        for (int m = 0; m < M; ++m) miss = miss || slot_empty(sv[m]);
        if (!__any(miss)) break;
        __builtin_amdgcn_s_sleep(1);
#pragma unroll
        for (int m = 0; m < M; ++m)
          if (slot_empty(sv[m]))
            sv[m] = __hip_atomic_load(Gl + (int64_t)(wave + m * kCoopWaves) * NBP, __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
        if (slot_empty(pv)) pv = __hip_atomic_load(Gl + (int64_t)nwg * NBP, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      double s = 0.0;
#pragma unroll
      for (int m = 0; m < M; ++m) s += sv[m];
      red2[wave][lane] = s;
      __syncthreads();
      if (wave == 0) {
        g = 0.0;
#pragma unroll
        for (int k = 0; k < kCoopWaves; ++k) g += red2[k][lane];
        pr = pv;
      }
      }
    } else if (flr) {
      double* G = base + (int64_t)(jj & 1) * rec;
      const unsigned target = (unsigned)jj + 1u;
      if (wave == 0) {
        if (lane < rem) {
          if (l2local) {
            __hip_atomic_store(G + (int64_t)w * NBP + lane, g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (w == 0)
              __hip_atomic_store(G + (int64_t)nwg * NBP + lane, pr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          } else {
            __hip_atomic_store(G + (int64_t)w * NBP + lane, g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (w == 0)
              __hip_atomic_store(G + (int64_t)nwg * NBP + lane, pr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          }
        }
        // the slot stores are complete (at the coherence point: the device's,
        // or this XCD's L2 for the shift-per-XCD form) before the flag goes out
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_waitcnt(kWaitVm0);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        if (lane == 0) {
          if (l2local) __hip_atomic_store(flw, target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          else __hip_atomic_store(flw, target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        for (unsigned spin = 0; spin < kCoopSpinCap; ++spin) {
          const unsigned f =
              lane < nwg ? __hip_atomic_load(flr + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : target;
          if (__all(f >= target)) break;
          __builtin_amdgcn_s_sleep(1);
        }
      }
      __syncthreads();
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      // this wave's share of the slots (workgroups wave, wave + 8, ...), all
      // loads in flight together: out-of-range entries read slot 0 and are masked
      double sv[kCoopMaxWg / kCoopWaves];
      const int cl = lane < rem ? lane : 0;
#pragma unroll
      for (int m = 0; m < kCoopMaxWg / kCoopWaves; ++m) {
        const int k = wave + m * kCoopWaves;
        sv[m] = __hip_atomic_load(G + (int64_t)(k < nwg ? k : 0) * NBP + cl, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
      }
      if (wave == 0) pr = __hip_atomic_load(G + (int64_t)nwg * NBP + cl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      double s = 0.0;
#pragma unroll
      for (int m = 0; m < kCoopMaxWg / kCoopWaves; ++m)
        if (wave + m * kCoopWaves < nwg) s += sv[m];
      red[wave][lane] = s;  // wave 0 read red before the barrier above
      __syncthreads();
      if (wave == 0) {
        g = 0.0;
#pragma unroll
        for (int k = 0; k < kCoopWaves; ++k) g += red[k][lane];
        if (lane >= rem) pr = 0.0;
      }
    }
    if (wave == 0) {
      // the reflector (dlarfg) from lane 0's r_0 and g_0, in every lane
      const double alpha = __shfl(pr, 0, 64), xn2 = __shfl(g, 0, 64);
      double beta = alpha, tv = 0.0, scal = 0.0;
      if (xn2 > 0.0) {
        beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
        tv = (beta - alpha) / beta;
        scal = 1.0 / (alpha - beta);
      }
      // w_c = tau (r_c + scal g_c): the update of column c is v_i w_c; zero past the panel
      if (lane < NV) wv[lane] = (lane >= 1 && lane < rem) ? tv * (pr + scal * g) : 0.0;
      if (lane == 0) {
        sc[0] = beta;
        sc[1] = scal;
        if (w == 0) tau[(int64_t)q * strideTau + j] = (T)tv;
      }
    }
    __syncthreads();
    if (live) {
      const double vi = (i == j) ? 1.0 : v[0] * sc[1];
      v[0] = (i == j) ? sc[0] : vi;
#pragma unroll
      for (int c = 1; c < NV; ++c) v[c] -= vi * wv[c];
    }
    // column jj is final in every row (rows above j were finished by earlier reflectors)
    if (valid) *out = (T)v[0];
    out += lds;
#pragma unroll
    for (int c = 0; c + 1 < NV; ++c) v[c] = v[c + 1];
    v[NV - 1] = 0.0;
    // No barrier here: red and prj are written next only after every thread
    // has passed this column's last barrier (wave 0 read them before it), wv
    // and sc only after the next column's first barrier.
  };
  int jj = 0;
  for (; jj < jb && jb - jj > 32; ++jj) column(std::integral_constant<int, 64>{}, jj);
  for (; jj < jb && jb - jj > 16; ++jj) column(std::integral_constant<int, 32>{}, jj);
  for (; jj < jb; ++jj) column(std::integral_constant<int, 16>{}, jj);
}

// Whole-panel factorisation in one workgroup for stacks of at most
// kQrPanelWgRows rows (the small matrices of the batched path): one thread per
// row with the row's NC panel values in registers, shifted so that the
// current column is always v[0] (no register is indexed by the loop counter);
// every step stores the finished column and shifts.  For each column one
// workgroup-wide reduction gives the Gram row g_c = sum_{i>j} S[i][j] S[i][c]
// (a wave butterfly, then the waves summed in FP64 in a fixed order).
// NC = 64 for FP32 panels, 32 for FP64 panels (register budget).
constexpr int kPanelWgRows = kQrPanelWgRows;

template <typename T, int NC>
__global__ __launch_bounds__(kPanelWgRows) void qr_panel_wg_kernel(T* S, int64_t lds, int64_t strideS, int64_t n,
                                                                   int64_t p, int jb, T* tau, int64_t strideTau) {
  __shared__ double red[kPanelWgRows / 64][NC];
  __shared__ double rr[NC];
  __shared__ T wv[NC];
  __shared__ T sc[2];
  const int q = blockIdx.y;
  T* Sq = S + q * strideS;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6, nwave = blockDim.x >> 6;
  const int64_t mrows = n + jb;
  const int64_t i = p + t;
  const bool mine = t < mrows;
  // columns >= jb are zero in registers and get zero updates throughout
  T v[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) v[c] = (mine && c < jb) ? Sq[i + (p + c) * lds] : (T)0;
  for (int jj = 0; jj < jb; ++jj) {
    const int64_t j = p + jj;
    const bool live = mine && i >= j && i <= n + j;
    if (live && i == j) {
#pragma unroll
      for (int c = 0; c < NC; ++c) rr[c] = (double)v[c];
    }
    const T x = (live && i > j) ? v[0] : (T)0;
    // lanes c >= 1 of the sum: the Gram row (columns jj + c); lane 0: the
    // squared norm below the diagonal
    T gs;
    if constexpr (NC == 64) {
      gs = wave_colsum64_swap<T>(v, x, lane);
    } else {
      T g[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) g[c] = x * v[c];
      gs = wave_colsum32_swap<T>(g, lane);
    }
    if (lane < NC) red[wave][lane] = (double)gs;
    __syncthreads();
    if (wave == 0) {
      double gc = 0.0;
      if (lane < NC)
        for (int w = 0; w < nwave; ++w) gc += red[w][lane];
      const double xn2 = __shfl(gc, 0, 64);
      const double alpha = rr[0];
      double beta = alpha, tv = 0.0, scal = 0.0;
      if (xn2 > 0.0) {
        beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
        tv = (beta - alpha) / beta;
        scal = 1.0 / (alpha - beta);
      }
      // w_c = tau v^T S[:, jj + c] for c >= 1, zero elsewhere
      if (lane < NC) wv[lane] = (lane >= 1 && lane < jb - jj) ? (T)(tv * (rr[lane] + scal * gc)) : (T)0;
      if (lane == 0) {
        sc[0] = (T)beta;
        sc[1] = (T)scal;
        tau[q * strideTau + j] = (T)tv;
      }
    }
    __syncthreads();
    if (live) {
      const T vi = (i == j) ? (T)1 : x * sc[1];
      v[0] = (i == j) ? sc[0] : vi;
#pragma unroll
      for (int c = 1; c < NC; ++c) v[c] = fma(-vi, wv[c], v[c]);
    }
    if (mine) Sq[i + j * lds] = v[0];
#pragma unroll
    for (int c = 0; c + 1 < NC; ++c) v[c] = v[c + 1];
    v[NC - 1] = (T)0;
  }
}

// V (mrows x jb, rows [p, r1_end) then [n, r2_end)) with unit diagonal and
// explicit zeros, from the panel at (p, p).
template <typename T>
__global__ __launch_bounds__(kT) void qr_extract_v_kernel(const T* S, int64_t lds, int64_t strideS, T* V, int64_t ldv,
                                                          int64_t strideV, int64_t p, int jb, int64_t n,
                                                          int64_t r1_end, int64_t r2_end) {
  const int q = blockIdx.y;
  const T* Sq = S + q * strideS;
  T* Vq = V + q * strideV;
  const int64_t len1 = r1_end - p, mrows = len1 + (r2_end - n);
  const int64_t total = mrows * jb;
  for (int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x; e < total; e += (int64_t)gridDim.x * kT) {
    const int64_t i = e % mrows;
    const int c = (int)(e / mrows);
    const int64_t a = i < len1 ? p + i : n + (i - len1), pc = p + c;
    T v = (T)0;
    if (a == pc) v = (T)1;
    else if (a > pc && (a < r1_end || a <= n + pc)) v = Sq[a + pc * lds];
    Vq[i + (int64_t)c * ldv] = v;
  }
}

// T (jb x jb upper) of the block reflector from G = V^T V and tau (dlarft,
// forward columnwise), in the recursive form: the four 16-column diagonal
// blocks by the column recurrence T[0:i, i] = -tau_i T[0:i, 0:i] G[0:i, i],
// all four at once, then pairs and halves joined by
// T12 = -T1 (V1^T V2) T2 = -T1 G12 T2: 38 barriers instead of 128.  Columns
// from jb to 63 are padded with tau = 0 and G = 0, which leaves them zero.
template <typename T>
__global__ __launch_bounds__(kT) void qr_form_t_kernel(const T* G, int64_t ldg, int64_t strideG, const T* tau,
                                                       int64_t strideTau, int64_t p, int jb, T* Tm, int64_t ldt,
                                                       int64_t strideT) {
  static_assert(NBP == 64 && kT == 256, "four 16-column blocks of 64 threads");
  __shared__ double g[NBP][NBP + 1];
  __shared__ double tt[NBP][NBP + 1];
  __shared__ double xx[NBP][NBP + 1];
  __shared__ double ts[NBP];
  const int q = blockIdx.x;
  const T* Gq = G + q * strideG;
  const T* tq = tau + q * strideTau;
  T* Tq = Tm + q * strideT;
  const int t = threadIdx.x;
  for (int e = t; e < NBP * NBP; e += kT) {
    const int r = e % NBP, c = e / NBP;
    g[r][c] = (r < jb && c < jb) ? (double)Gq[r + c * ldg] : 0.0;
    tt[r][c] = 0.0;
  }
  if (t < NBP) ts[t] = t < jb ? (double)tq[p + t] : 0.0;
  __syncthreads();
  {  // the diagonal blocks: block bk = t / 64, row rr of it, each dot split over four lanes
    const int bk = t >> 6, rr = (t >> 2) & 15, part = t & 3, o = bk * 16;
    for (int i = 0; i < 16; ++i) {
      const int col = o + i;
      double s = 0.0;
      if (rr < i)
        for (int k = rr + part; k < i; k += 4) s += tt[o + rr][o + k] * g[o + k][col];
      s += __shfl_xor(s, 1, 64);
      s += __shfl_xor(s, 2, 64);
      __syncthreads();
      if (part == 0 && rr < i) tt[o + rr][col] = -ts[col] * s;
      if (part == 0 && rr == i) tt[col][col] = ts[col];
      __syncthreads();
    }
  }
  {  // pairs (0, 1) and (2, 3): X = G12 T2, then T12 = -T1 X (16 x 16 each)
    const int r = t >> 4, c = t & 15;
#pragma unroll
    for (int P = 0; P < 2; ++P) {
      const int oa = 32 * P, ob = oa + 16;
      double x = 0.0;
      for (int k = 0; k <= c; ++k) x += g[oa + r][ob + k] * tt[ob + k][ob + c];
      xx[oa + r][ob + c] = x;
    }
    __syncthreads();
#pragma unroll
    for (int P = 0; P < 2; ++P) {
      const int oa = 32 * P, ob = oa + 16;
      double y = 0.0;
      for (int k = r; k < 16; ++k) y += tt[oa + r][oa + k] * xx[oa + k][ob + c];
      tt[oa + r][ob + c] = -y;
    }
    __syncthreads();
  }
  {  // halves: X = G[0:32, 32:64] T2, then T[0:32, 32:64] = -T1 X (32 x 32)
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const int e = t + kT * m, r = e >> 5, c = e & 31;
      double x = 0.0;
      for (int k = 0; k <= c; ++k) x += g[r][32 + k] * tt[32 + k][32 + c];
      xx[r][32 + c] = x;
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const int e = t + kT * m, r = e >> 5, c = e & 31;
      double y = 0.0;
      for (int k = r; k < 32; ++k) y += tt[r][k] * xx[k][32 + c];
      tt[r][32 + c] = -y;
    }
    __syncthreads();
  }
  for (int e = t; e < jb * jb; e += kT) {
    const int r = e % jb, c = e / jb;
    Tq[r + c * ldt] = (T)tt[r][c];
  }
}

template <typename T>
__global__ __launch_bounds__(kT) void identity_top_kernel(T* Q, int64_t ldq, int64_t strideQ, int64_t m, int64_t n) {
  T* Qq = Q + blockIdx.y * strideQ;
  const int64_t total = m * n;
  for (int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x; e < total; e += (int64_t)gridDim.x * kT) {
    const int64_t i = e % m, j = e / m;
    Qq[i + j * ldq] = (i == j) ? (T)1 : (T)0;
  }
}

inline unsigned grid_for(int64_t total) {
  int64_t g = (total + kT - 1) / kT;
  if (g > 8192) g = 8192;
  if (g < 1) g = 1;
  return (unsigned)g;
}

}  // namespace

template <typename T>
int k_qr_panel_step(T* S, int64_t lds, int64_t strideS, int64_t n, int64_t p, int jb, int64_t r1_end, int64_t r2_end,
                    int64_t j, int apply_prev, int reduce_cur, double* part, int nwg, T* tau, int64_t strideTau,
                    int batch, hipStream_t s) {
  if (jb > NBP || jb <= 0 || r1_end < p + jb || r1_end > n || r2_end < n || r2_end > n + p + jb ||
      (int64_t)nwg * kT < (r1_end - p) + (r2_end - n)) {
    set_error("qr panel: width or row ranges out of range");
    return kErrArg;
  }
  hipLaunchKernelGGL((qr_panel_kernel<T>), dim3(nwg, batch), dim3(kT), 0, s, S, lds, strideS, n, p, jb, r1_end, r2_end,
                     j, apply_prev, reduce_cur, part, nwg, tau, strideTau);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

// Workgroups of 256 threads (one wave per SIMD) where the launch leaves CUs
// idle, so that each wave issues its column sums and updates alone; 512
// threads (two waves per SIMD, half as many workgroups to exchange with)
// where the grid would fill the device anyway.
template <typename T, int TT>
int coop_cap(int dev) {
  static int cap[64] = {0};
  if (dev < 0 || dev >= 64) return -1;
  if (cap[dev] == 0) {
    // -1 (no cooperative launch) when the device cannot tell
    int per_cu = 0, cus = 0, coop = 0;
    const bool ok =
        hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)qr_panel_coop_kernel<T, TT>, TT, 0) ==
            hipSuccess;
    cap[dev] = ok && coop && per_cu * cus > 0 ? per_cu * cus : -1;
  }
  return cap[dev];
}

template <typename T, int TT>
int coop_launch(T* S, int64_t lds, int64_t strideS, int64_t n, int64_t p, int jb, int64_t r1_end, int64_t r2_end,
                double* part, int64_t part_doubles, T* tau, int64_t strideTau, int batch, int nwg, int cap,
                hipStream_t s) {
  const int per = std::min(batch, cap / nwg);
  const int64_t rec = (int64_t)nwg * NBP + NBP;
  if constexpr (TT <= 256) {
    // one slot record per panel column and shift, all empty at the start
    if ((int64_t)per * NBP * rec > part_doubles) {
      set_error("qr panel (cooperative): partials buffer too small");
      return kErrArg;
    }
    for (int q0 = 0; q0 < batch; q0 += per) {
      const int nq = std::min(per, batch - q0);
      int nw = nwg;
      unsigned* noflags = nullptr;
      ZOLO_HIP(hipMemsetAsync(part, 0xFF, sizeof(double) * (size_t)(nq * NBP * rec), s));
      int xmap = 0, nqa = nq;
      void* args[] = {&S, &lds, &strideS, &n, &p, &jb, &r1_end, &r2_end, &part, &noflags, &nw, &tau, &strideTau,
                      (void*)&q0, &xmap, &nqa};
      ZOLO_HIP(hipLaunchCooperativeKernel((const void*)qr_panel_coop_kernel<T, TT>, dim3(nwg, nq), dim3(TT), args,
                                          0, s));
    }
  } else {
    // two alternating slot records per shift, then the flags of each shift on
    // a 256-byte line of their own
    if ((int64_t)per * 2 * rec + (int64_t)per * kCoopFlagStride / 2 + 32 > part_doubles) {
      set_error("qr panel (cooperative): partials buffer too small");
      return kErrArg;
    }
    unsigned* flags = reinterpret_cast<unsigned*>(part + ((int64_t)per * 2 * rec + 31) / 32 * 32);
    for (int q0 = 0; q0 < batch; q0 += per) {
      const int nq = std::min(per, batch - q0);
      int nw = nwg;
      ZOLO_HIP(hipMemsetAsync(flags, 0, sizeof(unsigned) * kCoopFlagStride * nq, s));
      int xmap = 0, nqa = nq;
      void* args[] = {&S, &lds, &strideS, &n, &p, &jb, &r1_end, &r2_end, &part, &flags, &nw, &tau, &strideTau,
                      (void*)&q0, &xmap, &nqa};
      ZOLO_HIP(hipLaunchCooperativeKernel((const void*)qr_panel_coop_kernel<T, TT>, dim3(nwg, nq), dim3(TT), args,
                                          0, s));
    }
  }
  return kOk;
}

// Process-wide: the shift-per-XCD panels take device-scope stores even when
// their workgroups share one XCD (set by qr_set_panel_device_stores).
int g_panel_device_stores = 0;

// One shift per XCD (at most 8 shifts, at most one XCD's 32 CUs each): 8
// workgroups launched per row block, every eighth on the same XCD under the
// round-robin dispatch (the kernel checks), slot polling with L2-local stores.
template <typename T, int TT>
int coop_launch_xcd(T* S, int64_t lds, int64_t strideS, int64_t n, int64_t p, int jb, int64_t r1_end, int64_t r2_end,
                    double* part, int64_t part_doubles, T* tau, int64_t strideTau, int batch, int nwg, hipStream_t s) {
  const int64_t rec = (int64_t)nwg * NBP + NBP;
  const int64_t xid_doubles = (int64_t)batch * kCoopMaxWg / 2;
  unsigned* flags = nullptr;
  if constexpr (TT <= 256) {
    // slot records of every column, then the XCC words: all bits set (empty)
    const int64_t slots = (int64_t)batch * NBP * rec;
    if (slots + xid_doubles > part_doubles) {
      set_error("qr panel (cooperative): partials buffer too small");
      return kErrArg;
    }
    ZOLO_HIP(hipMemsetAsync(part, 0xFF, sizeof(double) * (size_t)(slots + xid_doubles), s));
  } else {
    // two alternating slot records, the XCC words (all bits set), the flags (zero)
    const int64_t slots = (int64_t)batch * 2 * rec, flag_doubles = (int64_t)batch * kCoopFlagStride / 2;
    if (slots + xid_doubles + flag_doubles > part_doubles) {
      set_error("qr panel (cooperative): partials buffer too small");
      return kErrArg;
    }
    ZOLO_HIP(hipMemsetAsync(part + slots, 0xFF, sizeof(double) * (size_t)xid_doubles, s));
    flags = reinterpret_cast<unsigned*>(part + slots + xid_doubles);
    ZOLO_HIP(hipMemsetAsync(flags, 0, sizeof(unsigned) * (size_t)batch * kCoopFlagStride, s));
  }
  // xmap 1: L2-local stores when the placement check passes; 2: device-scope
  // stores regardless (the form the check falls back to)
  int nw = nwg, q0 = 0, xmap = g_panel_device_stores ? 2 : 1, nq = batch;
  void* args[] = {&S, &lds, &strideS, &n, &p, &jb, &r1_end, &r2_end, &part, &flags, &nw, &tau, &strideTau,
                  (void*)&q0, &xmap, &nq};
  ZOLO_HIP(hipLaunchCooperativeKernel((const void*)qr_panel_coop_kernel<T, TT>, dim3(8 * nwg), dim3(TT), args, 0, s));
  return kOk;
}

template <typename T>
int k_qr_panel_coop(T* S, int64_t lds, int64_t strideS, int64_t n, int64_t p, int jb, int64_t r1_end, int64_t r2_end,
                    double* part, int64_t part_doubles, T* tau, int64_t strideTau, int batch, hipStream_t s,
                    bool* done) {
  *done = false;
  if (jb > NBP || jb <= 0 || r1_end < p + jb || r1_end > n || r2_end < n || r2_end > n + p + jb) {
    set_error("qr panel (cooperative): width or row ranges out of range");
    return kErrArg;
  }
  const int64_t rows = (r1_end - p) + (r2_end - n);
  if (rows <= 512) {
    // one workgroup per shift: no barrier between workgroups, an ordinary
    // launch (also inside a graph capture)
    if ((int64_t)batch * 2 * (2 * NBP) > part_doubles) {
      set_error("qr panel (one workgroup): partials buffer too small");
      return kErrArg;
    }
    int one = 1;
    unsigned* noflags = nullptr;
    int q0 = 0, xmap = 0;
    if (rows <= 256)
      hipLaunchKernelGGL((qr_panel_coop_kernel<T, 256>), dim3(1, batch), dim3(256), 0, s, S, lds, strideS, n, p, jb,
                         r1_end, r2_end, part, noflags, one, tau, strideTau, q0, xmap, batch);
    else
      hipLaunchKernelGGL((qr_panel_coop_kernel<T, 512>), dim3(1, batch), dim3(512), 0, s, S, lds, strideS, n, p, jb,
                         r1_end, r2_end, part, noflags, one, tau, strideTau, q0, xmap, batch);
    ZOLO_HIP(hipGetLastError());
    *done = true;
    return kOk;
  }
  {  // cooperative launches cannot be captured into a graph: take the per-column kernels there
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    ZOLO_HIP(hipStreamIsCapturing(s, &cs));
    if (cs != hipStreamCaptureStatusNone) return kOk;
  }
  int dev = 0, cus = 0;
  ZOLO_HIP(hipGetDevice(&dev));
  ZOLO_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int nwg256 = (int)((rows + 255) / 256), nwg512 = (int)((rows + 511) / 512);
  const int cap256 = coop_cap<T, 256>(dev), cap512 = coop_cap<T, 512>(dev);
  constexpr int kXcdCus = 32;  // CUs of one XCD: workgroups per shift in the shift-per-XCD form
  if (batch <= 8 && nwg256 <= kXcdCus && cap256 >= 8 * nwg256) {
    ZOLO_CHECK(coop_launch_xcd<T, 256>(S, lds, strideS, n, p, jb, r1_end, r2_end, part, part_doubles, tau, strideTau,
                                       batch, nwg256, s));
  } else if (batch >= 2 && batch <= 8 && nwg512 <= kXcdCus && cap512 >= 8 * nwg512) {
    // (one shift over 8192 rows: 64 CUs with one wave per SIMD beat one XCD's
    // 32 with two, 248-258 against 275-295 us per panel)
    ZOLO_CHECK(coop_launch_xcd<T, 512>(S, lds, strideS, n, p, jb, r1_end, r2_end, part, part_doubles, tau, strideTau,
                                       batch, nwg512, s));
  } else if (nwg256 <= kCoopMaxWg && (int64_t)batch * nwg256 <= cus && cap256 >= nwg256) {
    ZOLO_CHECK(coop_launch<T, 256>(S, lds, strideS, n, p, jb, r1_end, r2_end, part, part_doubles, tau, strideTau,
                                   batch, nwg256, cap256, s));
  } else if (nwg512 <= kCoopMaxWg && cap512 >= nwg512) {
    ZOLO_CHECK(coop_launch<T, 512>(S, lds, strideS, n, p, jb, r1_end, r2_end, part, part_doubles, tau, strideTau,
                                   batch, nwg512, cap512, s));
  } else {
    // too large for one resident grid or for the partial sums: the caller takes the per-column kernels
    return kOk;
  }
  *done = true;
  return kOk;
}

template <typename T>
int k_qr_panel_wg(T* S, int64_t lds, int64_t strideS, int64_t n, int64_t p, int jb, T* tau, int64_t strideTau,
                  int batch, hipStream_t s) {
  constexpr int NC = sizeof(T) == 4 ? 64 : 32;
  const int64_t mrows = n + jb;
  if (jb > NC || jb <= 0 || mrows > kPanelWgRows) {
    set_error("qr panel (one workgroup): size out of range");
    return kErrArg;
  }
  const unsigned threads = (unsigned)((mrows + 63) / 64 * 64);
  hipLaunchKernelGGL((qr_panel_wg_kernel<T, NC>), dim3(1, batch), dim3(threads), 0, s, S, lds, strideS, n, p, jb, tau,
                     strideTau);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

template <typename T>
int k_qr_extract_v(const T* S, int64_t lds, int64_t strideS, T* V, int64_t ldv, int64_t strideV, int64_t p, int jb,
                   int64_t n, int64_t r1_end, int64_t r2_end, int batch, hipStream_t s) {
  const int64_t mrows = (r1_end - p) + (r2_end - n);
  if (r1_end < p + jb || r1_end > n || r2_end < n || r2_end > n + p + jb || ldv < mrows) {
    set_error("qr extract v: row ranges out of range");
    return kErrArg;
  }
  hipLaunchKernelGGL((qr_extract_v_kernel<T>), dim3(grid_for(mrows * jb), batch), dim3(kT), 0, s, S, lds, strideS, V,
                     ldv, strideV, p, jb, n, r1_end, r2_end);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

template <typename T>
int k_qr_form_t(const T* G, int64_t ldg, int64_t strideG, const T* tau, int64_t strideTau, int64_t p, int jb, T* Tm,
                int64_t ldt, int64_t strideT, int batch, hipStream_t s) {
  hipLaunchKernelGGL((qr_form_t_kernel<T>), dim3(batch), dim3(kT), 0, s, G, ldg, strideG, tau, strideTau, p, jb, Tm,
                     ldt, strideT);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

template <typename T>
int k_set_identity_top(T* Q, int64_t ldq, int64_t strideQ, int64_t m, int64_t n, int batch, hipStream_t s) {
  hipLaunchKernelGGL((identity_top_kernel<T>), dim3(grid_for(m * n), batch), dim3(kT), 0, s, Q, ldq, strideQ, m, n);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

void qr_set_panel_device_stores(int on) { g_panel_device_stores = on ? 1 : 0; }
int qr_panel_device_stores() { return g_panel_device_stores; }

#define ZOLO_QR_INST(T)                                                                                            \
  template int k_qr_panel_step<T>(T*, int64_t, int64_t, int64_t, int64_t, int, int64_t, int64_t, int64_t, int, int, \
                                  double*, int, T*, int64_t, int, hipStream_t);                                    \
  template int k_qr_extract_v<T>(const T*, int64_t, int64_t, T*, int64_t, int64_t, int64_t, int, int64_t, int64_t,  \
                                 int64_t, int, hipStream_t);                                                       \
  template int k_qr_form_t<T>(const T*, int64_t, int64_t, const T*, int64_t, int64_t, int, T*, int64_t, int64_t,    \
                              int, hipStream_t);                                                                   \
  template int k_set_identity_top<T>(T*, int64_t, int64_t, int64_t, int64_t, int, hipStream_t);                \
  template int k_qr_panel_wg<T>(T*, int64_t, int64_t, int64_t, int64_t, int, T*, int64_t, int, hipStream_t);      \
  template int k_qr_panel_coop<T>(T*, int64_t, int64_t, int64_t, int64_t, int, int64_t, int64_t, double*, int64_t, \
                                  T*, int64_t, int, hipStream_t, bool*);
ZOLO_QR_INST(double)
ZOLO_QR_INST(float)

}  // namespace zolo
