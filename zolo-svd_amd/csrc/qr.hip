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
// in one launch with one barrier per column among the workgroups of a shift
// (qr_panel_coop_kernel); and, where neither applies (graph capture of a
// multi-workgroup panel, grids too large to be resident), one launch per
// column (qr_panel_kernel): launch j applies the reflector of column j-1 to
// the rest of the panel and, in the same pass over the rows each thread owns,
// reduces the Gram row g_c = sum_i S[i][j] S[i][c] that the reflector of
// column j needs (its norm is g_j).  Partial sums go to per-workgroup slots
// and are summed in a fixed order, so the result does not depend on
// scheduling.  The trailing update and the explicit Q use the compact WY form
// (I - V T V^T) through the MFMA GEMM.
#include <algorithm>

#include "grid_sync.hpp"
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
// by the loop counter), and stores each column when it is finished.  Per
// column: the Gram row g_c = sum_{i>j} S[i][j] S[i][j+c] and the pivot row go
// to a per-workgroup slot, one barrier among the workgroups of the shift, then
// every workgroup sums the slots in the same fixed order, forms the reflector
// (dlarfg) and applies it to its rows.  The slots alternate between two
// buffers, so the next column's partials never overwrite ones still being
// read.  Shifts q0 .. q0 + gridDim.y - 1; bar holds one zeroed counter per
// shift (kCoopBarStride words apart), or is null when nwg == 1 (no barrier
// between workgroups).
constexpr int kCoopT = 512;
constexpr int kCoopMaxSlots = 8;  // partial slots per thread when summing: nwg <= 8 * (kCoopT / 64)
constexpr int kCoopBarStride = 64;  // unsigned words between the barrier counters of two shifts

template <typename T>
__global__ __launch_bounds__(kCoopT) void qr_panel_coop_kernel(T* S, int64_t lds, int64_t strideS, int64_t n,
                                                               int64_t p, int jb, int64_t r1_end, int64_t r2_end,
                                                               double* part, unsigned* bar, int nwg, T* tau,
                                                               int64_t strideTau, int q0) {
  __shared__ double red[kCoopT / 64][NBP];
  __shared__ double wv[NBP], prj[NBP];
  __shared__ double sc[3];
  const int qi = blockIdx.y, q = q0 + qi;
  T* Sq = S + (int64_t)q * strideS;
  const int t = threadIdx.x, w = blockIdx.x, lane = t & 63, wave = t >> 6;
  const int64_t l = (int64_t)w * kCoopT + t, len1 = r1_end - p;
  const int64_t i = l < len1 ? p + l : n + (l - len1);
  const bool valid = l < len1 + (r2_end - n);
  const int64_t rec = (int64_t)nwg * NBP + NBP;
  double* base = part + (int64_t)qi * 2 * rec;
  // v[c]: the row's value in panel column jj + c; zero for jj + c >= jb
  double v[NBP];
#pragma unroll
  for (int c = 0; c < NBP; ++c) v[c] = (valid && c < jb) ? (double)Sq[i + (p + c) * lds] : 0.0;
  T* out = Sq + i + p * lds;
  for (int jj = 0; jj < jb; ++jj) {
    const int64_t j = p + jj;
    const int rem = jb - jj;  // columns jj .. jb - 1 are v[0 .. rem)
    double* G = base + (int64_t)(jj & 1) * rec;
    // bottom rows beyond n + j take no part in the reflector of column j
    const bool live = valid && i >= j && (i < r1_end || i <= n + j);
    if (live && i == j) {
#pragma unroll
      for (int c = 0; c < NBP; ++c) prj[c] = v[c];
    }
    const double x = (live && i > j) ? v[0] : 0.0;
    // lane 0: the squared norm below the diagonal; lanes c >= 1: the Gram row
    red[wave][lane] = wave_colsum64_swap(v, x, lane);
    __syncthreads();
    // wave 0 publishes the slots (the barrier's release is issued by its thread 0);
    // row j (the pivot row) lies in the top range, in workgroup (j - p) / kCoopT
    if (t < rem) {
      double sum = 0.0;
#pragma unroll
      for (int k = 0; k < kCoopT / 64; ++k) sum += red[k][t];
      G[(int64_t)w * NBP + t] = sum;
      if ((j - p) / kCoopT == w) G[(int64_t)nwg * NBP + t] = prj[t];
    }
    if (bar) grid_barrier(bar + (int64_t)qi * kCoopBarStride, (unsigned)(jj + 1) * (unsigned)nwg);
    else __syncthreads();  // one workgroup per shift
    // the pivot row rides along with the first group's partial loads (wave 0:
    // lane c holds r_c)
    double pr = 0.0;
    {  // sum the nwg partials of each column: 8 groups of 64 threads, all loads independent
      const int c = t & 63, g = t >> 6;
      double pv[kCoopMaxSlots];
#pragma unroll
      for (int m = 0; m < kCoopMaxSlots; ++m) {
        const int k = g + m * (kCoopT / 64);
        pv[m] = (c < rem && k < nwg) ? G[(int64_t)k * NBP + c] : 0.0;
      }
      if (g == 0 && c < rem) pr = G[(int64_t)nwg * NBP + c];
      double sum = 0.0;
#pragma unroll
      for (int m = 0; m < kCoopMaxSlots; ++m) sum += pv[m];
      red[g][c] = sum;
    }
    __syncthreads();
    if (wave == 0) {  // lane c: g_c, then the reflector (dlarfg) from lane 0's r_0 and g_0
      double gsum = 0.0;
      if (t < rem) {
#pragma unroll
        for (int g = 0; g < kCoopT / 64; ++g) gsum += red[g][t];
      }
      const double alpha = __shfl(pr, 0, 64), xn2 = __shfl(gsum, 0, 64);
      double beta = alpha, tv = 0.0, scal = 0.0;
      if (xn2 > 0.0) {
        beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
        tv = (beta - alpha) / beta;
        scal = 1.0 / (alpha - beta);
      }
      // w_c = tau (r_c + scal g_c): the update of column c is v_i w_c; zero past the panel
      wv[t] = (t >= 1 && t < rem) ? tv * (pr + scal * gsum) : 0.0;
      if (t == 0) {
        sc[0] = beta; sc[1] = tv; sc[2] = scal;
        if (w == 0) tau[(int64_t)q * strideTau + j] = (T)tv;
      }
    }
    __syncthreads();
    if (live) {
      const double vi = (i == j) ? 1.0 : v[0] * sc[2];
      v[0] = (i == j) ? sc[0] : vi;
#pragma unroll
      for (int c = 1; c < NBP; ++c) v[c] -= vi * wv[c];
    }
    // column jj is final in every row (rows above j were finished by earlier reflectors)
    if (valid) *out = (T)v[0];
    out += lds;
#pragma unroll
    for (int c = 0; c + 1 < NBP; ++c) v[c] = v[c + 1];
    v[NBP - 1] = 0.0;
    // No barrier here: the next column writes prj and red only after this
    // column's last barrier, which every reader of them has passed, and wv/sc
    // only after the next grid barrier.
  }
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

// T (jb x jb upper) of the block reflector from G = V^T V and tau (dlarft, forward columnwise).
template <typename T>
__global__ __launch_bounds__(kT) void qr_form_t_kernel(const T* G, int64_t ldg, int64_t strideG, const T* tau,
                                                       int64_t strideTau, int64_t p, int jb, T* Tm, int64_t ldt,
                                                       int64_t strideT) {
  __shared__ double g[NBP][NBP + 1];
  __shared__ double tt[NBP][NBP + 1];
  __shared__ double ts[NBP];
  const int q = blockIdx.x;
  const T* Gq = G + q * strideG;
  const T* tq = tau + q * strideTau;
  T* Tq = Tm + q * strideT;
  const int t = threadIdx.x;
  for (int e = t; e < jb * jb; e += kT) {
    const int r = e % jb, c = e / jb;
    g[r][c] = (double)Gq[r + c * ldg];
    tt[r][c] = 0.0;
  }
  if (t < jb) ts[t] = (double)tq[p + t];  // all of tau up front: no global load inside the column loop
  __syncthreads();
  // column i: T[0:i, i] = -tau_i T[0:i, 0:i] G[0:i, i]; each row's dot is
  // split over four lanes (k = row + part, row + part + 4, ...) and combined by
  // two shuffles inside the group of four
  const int row = t >> 2, part = t & 3;
  for (int i = 0; i < jb; ++i) {
    const double ti = ts[i];
    double s = 0.0;
    if (row < i)
      for (int k = row + part; k < i; k += 4) s += tt[row][k] * g[k][i];
    s += __shfl_xor(s, 1, 64);
    s += __shfl_xor(s, 2, 64);
    __syncthreads();
    if (part == 0 && row < i) tt[row][i] = -ti * s;
    if (t == 0) tt[i][i] = ti;
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

template <typename T>
int k_qr_panel_coop(T* S, int64_t lds, int64_t strideS, int64_t n, int64_t p, int jb, int64_t r1_end, int64_t r2_end,
                    double* part, int64_t part_doubles, T* tau, int64_t strideTau, int batch, hipStream_t s,
                    bool* done) {
  *done = false;
  if (jb > NBP || jb <= 0 || r1_end < p + jb || r1_end > n || r2_end < n || r2_end > n + p + jb) {
    set_error("qr panel (cooperative): width or row ranges out of range");
    return kErrArg;
  }
  const int64_t rows0 = (r1_end - p) + (r2_end - n);
  if (rows0 <= kCoopT) {
    // one workgroup per shift: no barrier between workgroups, an ordinary
    // launch (also inside a graph capture)
    if ((int64_t)batch * 2 * (2 * NBP) > part_doubles) {
      set_error("qr panel (one workgroup): partials buffer too small");
      return kErrArg;
    }
    int one = 1;
    unsigned* nobar = nullptr;
    int q0 = 0;
    hipLaunchKernelGGL((qr_panel_coop_kernel<T>), dim3(1, batch), dim3(kCoopT), 0, s, S, lds, strideS, n, p, jb,
                       r1_end, r2_end, part, nobar, one, tau, strideTau, q0);
    ZOLO_HIP(hipGetLastError());
    *done = true;
    return kOk;
  }
  {  // cooperative launches cannot be captured into a graph: take the per-column kernels there
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    ZOLO_HIP(hipStreamIsCapturing(s, &cs));
    if (cs != hipStreamCaptureStatusNone) return kOk;
  }
  // workgroups one launch may keep resident, per device
  static int cap[64] = {0};
  int dev = 0;
  ZOLO_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return kOk;
  if (cap[dev] == 0) {
    int per_cu = 0, cus = 0, coop = 0;
    ZOLO_HIP(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
    ZOLO_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    ZOLO_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)qr_panel_coop_kernel<T>, kCoopT, 0));
    cap[dev] = coop ? std::max(per_cu * cus, -1) : -1;
    if (cap[dev] == 0) cap[dev] = -1;
  }
  const int64_t rows = (r1_end - p) + (r2_end - n);
  const int nwg = (int)((rows + kCoopT - 1) / kCoopT);
  // too large for one resident grid or for the partial sums: the caller takes the per-column kernels
  if (cap[dev] < nwg || nwg > kCoopMaxSlots * (kCoopT / 64)) return kOk;
  const int per = std::min(batch, cap[dev] / nwg);
  const int64_t rec = (int64_t)nwg * NBP + NBP;
  // partial slots of the group, then one barrier counter per shift, each on a
  // 256-byte line of its own (counters sharing a line serialise the arrivals
  // of all shifts: 925 us instead of 615 us for 8 stacks of 16384 x 64)
  if ((int64_t)per * 2 * rec + (int64_t)per * kCoopBarStride / 2 + 32 > part_doubles) {
    set_error("qr panel (cooperative): partials buffer too small");
    return kErrArg;
  }
  unsigned* bar = reinterpret_cast<unsigned*>(part + ((int64_t)per * 2 * rec + 31) / 32 * 32);
  for (int q0 = 0; q0 < batch; q0 += per) {
    const int nq = std::min(per, batch - q0);
    int nw = nwg;
    ZOLO_HIP(hipMemsetAsync(bar, 0, sizeof(unsigned) * kCoopBarStride * nq, s));
    void* args[] = {&S, &lds, &strideS, &n, &p, &jb, &r1_end, &r2_end, &part, &bar, &nw, &tau, &strideTau,
                    (void*)&q0};
    ZOLO_HIP(hipLaunchCooperativeKernel((const void*)qr_panel_coop_kernel<T>, dim3(nwg, nq), dim3(kCoopT), args, 0,
                                        s));
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
