// Symmetric eigensolver of the Zolo-SVD tail (H = V diag(lambda) V^T, the
// role of ELPA_PDSYEVD2 at ZoloCHK1.f90:86), FP64:
//
//  1. Tridiagonalisation Q^T H Q = T (LAPACK dsytrd / dlatrd, lower): panels
//     of kTrdNb columns; per column a fused column update + Householder, one
//     matrix-vector product with the trailing matrix (the HBM-bound part,
//     read once per column), and the W column; per panel a rank-2k update of
//     the trailing matrix on the MFMA GEMM.
//  2. Tridiagonal divide and conquer (Cuppen; deflation as dlaed2, secular
//     equation by bisection on the offset from the nearest pole, eigenvectors
//     by the Gu-Eisenstat formula so they stay orthogonal), bottom-up over
//     levels with every merge of a level in the same launches; the eigenvector
//     update of each merge is an MFMA GEMM.
//  3. Back-transformation V = Q Z with the reflectors of step 1 in blocks
//     (compact WY, MFMA GEMMs).
#include <cfloat>

#include "eig.hpp"
#include "wave_sums.hpp"

namespace zolo {
namespace {

constexpr int kT = 256;

__device__ inline double block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0.0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r += sh[w];
  __syncthreads();
  return r;
}

inline unsigned blocks_for(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kT - 1) / kT); }

// ---------------------------------------------------------------------------
// 1. Tridiagonalisation.  A (n x n, ld) full symmetric; panel columns k0 ..;
//    W (n x kTrdNb, ld n) the dlatrd W of the panel; v (n) the current
//    reflector; sc[] device scalars; part[] per-workgroup partial sums.
// ---------------------------------------------------------------------------

// Column i (panel-local j): A[i:, i] -= A[i:, k0:i] W[i, 0:j]^T + W[i:, 0:j] A[i, k0:i]^T
// (the earlier reflectors of the panel), partial sums of squares of
// A[i+2:, i] for the reflector, and sc[0] = the updated A[i+1][i].
// fix: column j - 1 of W still lacks dlatrd's last correction W[:, j-1] +=
// alpha v (alpha = -tau[i-1]/2 w^T v from wpart, v = reflector i - 1); it is
// applied here to rows > i, which this launch owns, and to the row-i values
// it reads (row i of W is read by nothing else: the later columns read rows
// > i, the rank-2k update rows >= k0 + the panel width).
__global__ __launch_bounds__(kT) void trd_col_kernel(double* A, int64_t ld, int64_t n, int64_t k0, int64_t i,
                                                     double* W, double* part, double* sc, const double* v,
                                                     const double* wpart, int nwp, const double* tau, int fix) {
  __shared__ double arow[kTrdNb], wrow[kTrdNb];
  __shared__ double sh[4];
  __shared__ double s_alpha;
  const int j = (int)(i - k0);
  const int t = threadIdx.x;
  if (t == 0) {
    double a = 0.0;
    if (fix) {
      double s = 0.0;
      for (int k = 0; k < nwp; ++k) s += wpart[k];
      a = -0.5 * tau[i - 1] * s;
    }
    s_alpha = a;
  }
  __syncthreads();
  const double alpha = s_alpha;
  if (t < j) {
    arow[t] = A[i + (k0 + t) * ld];
    double wv = W[i + (int64_t)t * n];
    if (fix && t == j - 1) wv += alpha * v[i];
    wrow[t] = wv;
  }
  __syncthreads();
  double ss = 0.0;
  for (int64_t r = (int64_t)blockIdx.x * kT + t + i; r < n; r += (int64_t)gridDim.x * kT) {
    double a = A[r + i * ld];
    double wlast = 0.0;  // W[r][j-1] after the correction
    if (fix) {
      wlast = W[r + (int64_t)(j - 1) * n] + alpha * v[r];
      if (r > i) W[r + (int64_t)(j - 1) * n] = wlast;
    }
    for (int c = 0; c < j; ++c) {
      const double wrc = (fix && c == j - 1) ? wlast : W[r + (int64_t)c * n];
      a -= A[r + (k0 + c) * ld] * wrow[c] + wrc * arow[c];
    }
    A[r + i * ld] = a;
    if (r >= i + 2) ss += a * a;
    if (r == i + 1) sc[0] = a;
  }
  ss = block_sum(ss, sh);
  if (t == 0) part[blockIdx.x] = ss;
}

// Householder of A[i+1:, i] (dlarfg): tau[i], e[i] = beta; v = [1; A[i+2:, i] * scal],
// stored back into A with the unit head (used by the rest of the panel and the
// rank-2k update); partial dots of v with the panel columns of W and A.
__global__ __launch_bounds__(kT) void trd_house_kernel(double* A, int64_t ld, int64_t n, int64_t k0, int64_t i,
                                                       const double* W, const double* part, int np, const double* sc,
                                                       double* v, double* tau, double* e, double* dpart) {
  __shared__ double s_scal;
  __shared__ double dots[2 * kTrdNb][4];
  const int j = (int)(i - k0);
  const int t = threadIdx.x;
  if (t == 0) {
    double xn2 = 0.0;
    for (int k = 0; k < np; ++k) xn2 += part[k];
    const double alpha = sc[0];
    double beta = alpha, tv = 0.0, scal = 0.0;
    if (xn2 > 0.0) {
      beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
      tv = (beta - alpha) / beta;
      scal = 1.0 / (alpha - beta);
    }
    s_scal = scal;
    if (blockIdx.x == 0) {
      tau[i] = tv;
      e[i] = beta;
    }
  }
  __syncthreads();
  const double scal = s_scal;
  // dots of v with the panel columns of W (slot c < j) and A (slot kTrdNb + c):
  // each thread forms its rows' 64 products (all loads issued together), one
  // lane-swap butterfly gives every slot's wave sum
  static_assert(2 * kTrdNb == 64, "one 64-slot column sum");
  const int lane = t & 63, wave = t >> 6;
  double p[2 * kTrdNb];
#pragma unroll
  for (int c = 0; c < 2 * kTrdNb; ++c) p[c] = 0.0;
  for (int64_t r = (int64_t)blockIdx.x * kT + t + i + 1; r < n; r += (int64_t)gridDim.x * kT) {
    double x = 1.0;
    if (r > i + 1) x = A[r + i * ld] * scal;
    A[r + i * ld] = x;
    v[r] = x;
#pragma unroll
    for (int c = 0; c < kTrdNb; ++c)
      if (c < j) {
        p[c] += W[r + (int64_t)c * n] * x;
        p[kTrdNb + c] += A[r + (k0 + c) * ld] * x;
      }
  }
  for (int64_t r = (int64_t)blockIdx.x * kT + t; r <= i; r += (int64_t)gridDim.x * kT) v[r] = 0.0;
  dots[lane][wave] = wave_colsum64_swap(p, 1.0, lane);
  __syncthreads();
  if (t < 2 * kTrdNb) {
    const int c = t < kTrdNb ? t : t - kTrdNb;
    dpart[(int64_t)blockIdx.x * 2 * kTrdNb + t] =
        c < j ? (dots[t][0] + dots[t][1]) + (dots[t][2] + dots[t][3]) : 0.0;
  }
}

// y-partials of A22 v for the symmetric trailing matrix A22 (rows/cols >= i+1),
// reading only its lower triangle in 64 x 64 tiles: tile (R, C), R >= C,
// adds A_RC v_C to the rows of R (slot C) and, below the diagonal, A_RC^T v_R
// to the rows of C (slot R).  Every row r of block b thus gets exactly one
// partial per slot k < T: partial[k][r].  One workgroup per tile, the tile in LDS.
constexpr int kSymT = 64;
__global__ __launch_bounds__(kT) void trd_symv_kernel(const double* A, int64_t ld, int64_t n, int64_t i,
                                                      const double* v, double* partial) {
  __shared__ double tile[kSymT][kSymT + 1];
  __shared__ double vr[kSymT], vc[kSymT];
  __shared__ double red[4][kSymT];
  const int64_t r0 = i + 1;
  // tile index -> (R, C), R >= C, row-major over the lower triangle
  const int64_t tix = blockIdx.x;
  int64_t R = (int64_t)((sqrt(8.0 * (double)tix + 1.0) - 1.0) * 0.5);
  while ((R + 1) * (R + 2) / 2 <= tix) ++R;
  while (R * (R + 1) / 2 > tix) --R;
  const int64_t C = tix - R * (R + 1) / 2;
  const int64_t rb = r0 + R * kSymT, cb = r0 + C * kSymT;
  const int t = threadIdx.x;
  // load the tile (column-major in A: 64 contiguous rows per column)
  for (int e = t; e < kSymT * kSymT; e += kT) {
    const int rr = e % kSymT, cc = e / kSymT;
    const int64_t row = rb + rr, col = cb + cc;
    tile[rr][cc] = (row < n && col < n) ? A[row + col * ld] : 0.0;
  }
  if (t < kSymT) {
    vr[t] = (rb + t < n) ? v[rb + t] : 0.0;
    vc[t] = (cb + t < n) ? v[cb + t] : 0.0;
  }
  __syncthreads();
  const int lane = t & 63, q = t >> 6;  // 4 groups of 64: each sums 16 of the 64 terms
  {  // rows of R: y_r += sum_c tile[r][c] vc[c]
    double s0 = 0.0;
#pragma unroll
    for (int k = 0; k < kSymT / 4; ++k) {
      const int c = q * (kSymT / 4) + k;
      s0 += tile[lane][c] * vc[c];
    }
    red[q][lane] = s0;
  }
  __syncthreads();
  if (t < kSymT && rb + t < n) partial[C * n + rb + t] = (red[0][t] + red[1][t]) + (red[2][t] + red[3][t]);
  if (R == C) return;
  __syncthreads();
  {  // rows of C: y_c += sum_r tile[r][c] vr[r]
    double s1 = 0.0;
#pragma unroll
    for (int k = 0; k < kSymT / 4; ++k) {
      const int rr = q * (kSymT / 4) + k;
      s1 += tile[rr][lane] * vr[rr];
    }
    red[q][lane] = s1;
  }
  __syncthreads();
  if (t < kSymT && cb + t < n) partial[R * n + cb + t] = (red[0][t] + red[1][t]) + (red[2][t] + red[3][t]);
}

// w = tau (A22 v - Vp t_w - Wp t_a), t_w = Wp^T v, t_a = Vp^T v (from dpart);
// W[:, j] = w (rows >= i+1) and partial dots w^T v for the alpha correction.
__global__ __launch_bounds__(kT) void trd_w_kernel(const double* A, int64_t ld, int64_t n, int64_t k0, int64_t i,
                                                   double* W, const double* v, const double* partial, int nchunk,
                                                   const double* dpart, int ndp, const double* tau, double* wpart) {
  __shared__ double tw[kTrdNb], ta[kTrdNb];
  __shared__ double sh[4];
  const int j = (int)(i - k0);
  const int t = threadIdx.x;
  if (t < 2 * kTrdNb) {
    double s = 0.0;
    for (int k = 0; k < ndp; ++k) s += dpart[(int64_t)k * 2 * kTrdNb + t];
    if (t < kTrdNb) tw[t] = s;
    else ta[t - kTrdNb] = s;
  }
  __syncthreads();
  const double tv = tau[i];
  double acc = 0.0;
  for (int64_t r = (int64_t)blockIdx.x * kT + t + i + 1; r < n; r += (int64_t)gridDim.x * kT) {
    // four independent accumulators, so the loads are issued ahead of the adds
    double y0 = 0.0, y1 = 0.0, y2 = 0.0, y3 = 0.0;
    int k = 0;
    for (; k + 4 <= nchunk; k += 4) {
      y0 += partial[(int64_t)k * n + r];
      y1 += partial[(int64_t)(k + 1) * n + r];
      y2 += partial[(int64_t)(k + 2) * n + r];
      y3 += partial[(int64_t)(k + 3) * n + r];
    }
    for (; k < nchunk; ++k) y0 += partial[(int64_t)k * n + r];
    double z0 = 0.0, z1 = 0.0;
    int c = 0;
    for (; c + 2 <= j; c += 2) {
      z0 += A[r + (k0 + c) * ld] * tw[c] + W[r + (int64_t)c * n] * ta[c];
      z1 += A[r + (k0 + c + 1) * ld] * tw[c + 1] + W[r + (int64_t)(c + 1) * n] * ta[c + 1];
    }
    for (; c < j; ++c) z0 += A[r + (k0 + c) * ld] * tw[c] + W[r + (int64_t)c * n] * ta[c];
    const double y = ((y0 + y1) + (y2 + y3)) - (z0 + z1);
    const double w = tv * y;
    W[r + (int64_t)j * n] = w;
    acc += w * v[r];
  }
  for (int64_t r = (int64_t)blockIdx.x * kT + t; r <= i; r += (int64_t)gridDim.x * kT) W[r + (int64_t)j * n] = 0.0;
  acc = block_sum(acc, sh);
  if (t == 0) wpart[blockIdx.x] = acc;
}

// W[:, j] += alpha v, alpha = -tau/2 (w^T v) from the partial dots (dlatrd's last correction).
__global__ __launch_bounds__(kT) void trd_wfix_kernel(double* W, int64_t n, int j, const double* v,
                                                      const double* wpart, int nwp, const double* tau, int64_t i) {
  __shared__ double alpha;
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int k = 0; k < nwp; ++k) s += wpart[k];
    alpha = -0.5 * tau[i] * s;
  }
  __syncthreads();
  for (int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x + i + 1; r < n; r += (int64_t)gridDim.x * kT)
    W[r + (int64_t)j * n] += alpha * v[r];
}

// d[k] = A[k][k]; A[k+1][k] = 1 for the panel's reflectors (unit head for the
// rank-2k update and the back-transformation).
__global__ void trd_panel_end_kernel(double* A, int64_t ld, int64_t k0, int jb, int64_t n, double* d) {
  const int c = threadIdx.x;
  if (c < jb) {
    const int64_t k = k0 + c;
    d[k] = A[k + k * ld];
    if (k + 1 < n) A[(k + 1) + k * ld] = 1.0;
  }
}

// ---------------------------------------------------------------------------
// 2. Tridiagonal divide and conquer.
// ---------------------------------------------------------------------------
struct Merge {
  int off, s1, s2;
};

// Per merge (one workgroup): z, the merged order, deflation, and the
// compacted secular-equation data.
//   lam[off + c]   eigenvalues of the children (each child ascending)
//   Q              current eigenvector blocks (column c of the block at off)
// Outputs (indexed by off + position):
//   dK, zK, pK     non-deflated poles (ascending), weights, block column
//   dD, pD         deflated eigenvalues and block columns
//   rot            Givens rotations (col_a, col_b, c, s) to apply to Q
//   cnt[m]         (k, ndefl, nrot)
__global__ __launch_bounds__(1024) void dc_prepare_kernel(const Merge* merges, const double* e, const double* lam,
                                                          const double* Q, int64_t ldq, double* rho_out, double* dK,
                                                          double* zK, int* pK, double* dD, int* pD, double* rotcs,
                                                          int* rotab, int* cnt, double* zbuf, double* dbuf,
                                                          int* perm) {
  const Merge mg = merges[blockIdx.x];
  const int off = mg.off, s1 = mg.s1, s2 = mg.s2, s = s1 + s2;
  const int t = threadIdx.x;
  const double eb = e[off + s1 - 1];
  const double rho = 2.0 * fabs(eb);
  const double sg = eb >= 0.0 ? 1.0 : -1.0;
  // z (normalised by 1/sqrt(2)) in block column order
  for (int c = t; c < s; c += blockDim.x) {
    double z;
    if (c < s1) z = Q[(off + s1 - 1) + (int64_t)(off + c) * ldq];
    else z = sg * Q[(off + s1) + (int64_t)(off + c) * ldq];
    zbuf[off + c] = z * 0.70710678118654752440;
  }
  // merged ascending order of the two ascending children (rank merge)
  for (int c = t; c < s; c += blockDim.x) {
    const double x = lam[off + c];
    int lo, hi, rank;
    if (c < s1) {  // count right elements < x
      lo = 0; hi = s2;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (lam[off + s1 + mid] < x) lo = mid + 1; else hi = mid;
      }
      rank = c + lo;
    } else {  // count left elements <= x
      lo = 0; hi = s1;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (lam[off + mid] <= x) lo = mid + 1; else hi = mid;
      }
      rank = (c - s1) + lo;
    }
    perm[off + rank] = c;
    dbuf[off + rank] = x;
  }
  __syncthreads();
  if (t != 0) return;
  // deflation scan in ascending order (dlaed2): tol = 8 eps max(|d|max, rho |z|max)
  double dmax = 0.0, zmax = 0.0;
  for (int q = 0; q < s; ++q) {
    dmax = fmax(dmax, fabs(dbuf[off + q]));
    zmax = fmax(zmax, fabs(zbuf[off + q]));
  }
  const double tol = 8.0 * DBL_EPSILON * fmax(dmax, rho * zmax);
  int k = 0, nd = 0, nr = 0;
  int prev = -1;       // sorted position of the last non-deflated candidate
  double dprev = 0.0;  // its (possibly rotated) eigenvalue
  double zprev = 0.0;
  for (int q = 0; q < s; ++q) {
    const int col = perm[off + q];
    const double dq = dbuf[off + q];
    double zq = zbuf[off + col];
    if (rho * fabs(zq) <= tol) {  // negligible weight
      dD[off + nd] = dq;
      pD[off + nd] = col;
      ++nd;
      continue;
    }
    if (prev >= 0) {
      const double tau_ = hypot(zprev, zq);
      const double c = zq / tau_, sn = -zprev / tau_;
      if (fabs((dq - dprev) * c * sn) <= tol) {
        // rotate (prev, q): the prev direction loses its weight and deflates
        const int pc = perm[off + prev];
        rotab[2 * (off + nr)] = pc;
        rotab[2 * (off + nr) + 1] = col;
        rotcs[2 * (off + nr)] = c;
        rotcs[2 * (off + nr) + 1] = sn;
        ++nr;
        const double dp_new = dprev * c * c + dq * sn * sn;
        const double dq_new = dprev * sn * sn + dq * c * c;
        dD[off + nd] = dp_new;
        pD[off + nd] = pc;
        ++nd;
        prev = q;
        dprev = dq_new;
        zprev = tau_;
        continue;
      }
      // prev is final
      dK[off + k] = dprev;
      zK[off + k] = zprev;
      pK[off + k] = perm[off + prev];
      ++k;
    }
    prev = q;
    dprev = dq;
    zprev = zq;
  }
  if (prev >= 0) {
    dK[off + k] = dprev;
    zK[off + k] = zprev;
    pK[off + k] = perm[off + prev];
    ++k;
  }
  cnt[3 * blockIdx.x] = k;
  cnt[3 * blockIdx.x + 1] = nd;
  cnt[3 * blockIdx.x + 2] = nr;
  rho_out[blockIdx.x] = rho;
}

// Givens rotations of a merge on the block columns of Q (thread per row).
__global__ __launch_bounds__(kT) void dc_rotate_kernel(const Merge* merges, double* Q, int64_t ldq, const int* cnt,
                                                       const double* rotcs, const int* rotab) {
  const Merge mg = merges[blockIdx.y];
  const int off = mg.off, s = mg.s1 + mg.s2;
  const int nr = cnt[3 * blockIdx.y + 2];
  const int r = blockIdx.x * kT + threadIdx.x;
  if (r >= s || nr == 0) return;
  double* Qr = Q + (off + r);
  for (int q = 0; q < nr; ++q) {
    const int a = rotab[2 * (off + q)], b = rotab[2 * (off + q) + 1];
    const double c = rotcs[2 * (off + q)], sn = rotcs[2 * (off + q) + 1];
    const double qa = Qr[(int64_t)(off + a) * ldq], qb = Qr[(int64_t)(off + b) * ldq];
    // new directions: a' = c a + s b (weight 0), b' = -s a + c b (weight tau)
    Qr[(int64_t)(off + a) * ldq] = c * qa + sn * qb;
    Qr[(int64_t)(off + b) * ldq] = -sn * qa + c * qb;
  }
}

// f(tau) = 1 + rho sum z_i^2 / (delta_i - tau), delta_i = d_i - d_origin
// The secular function 1 + rho sum_i z_i^2 / ((d_i - d_org) - tau), summed
// by the 64 lanes of a wave (lane l takes terms l, l + 64, ...) in a fixed
// order; lane 0's total is broadcast, so every lane takes the same branch.
__device__ inline double secular_wave(const double* dK, const double* zK, int k, double dorg, double tau, double rho,
                                      int lane) {
  double s = 0.0;
  for (int i = lane; i < k; i += 64) s += rho * zK[i] * zK[i] / ((dK[i] - dorg) - tau);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  return 1.0 + __shfl(s, 0, 64);
}

// Roots of the secular equation by bisection on the offset from the nearer
// pole, one wave per root (four roots per workgroup).  Root j in
// (d_j, d_{j+1}), the last in (d_k, d_k + rho |z|^2).
constexpr int kRootsPerBlock = kT / 64;
__global__ __launch_bounds__(kT) void dc_roots_kernel(const Merge* merges, const int* cnt, const double* rho_in,
                                                      const double* dK_all, const double* zK_all, int* org,
                                                      double* tau_out) {
  const Merge mg = merges[blockIdx.y];
  const int off = mg.off;
  const int k = cnt[3 * blockIdx.y];
  const int lane = threadIdx.x & 63;
  const int jr = blockIdx.x * kRootsPerBlock + (threadIdx.x >> 6);
  if (jr >= k) return;  // the whole wave (no workgroup barrier in this kernel)
  const double* dK = dK_all + off;
  const double* zK = zK_all + off;
  const double rho = rho_in[blockIdx.y];
  int o;
  double lo, hi;
  if (jr < k - 1) {
    const double gap = dK[jr + 1] - dK[jr];
    const double fm = secular_wave(dK, zK, k, dK[jr], 0.5 * gap, rho, lane);
    if (fm >= 0.0) {  // root in the left half: offset from d_j
      o = jr; lo = 0.0; hi = 0.5 * gap;
    } else {          // right half: offset from d_{j+1} (negative)
      o = jr + 1; lo = -0.5 * gap; hi = 0.0;
    }
  } else {
    o = jr; lo = 0.0;
    double zz = 0.0;
    for (int i = lane; i < k; i += 64) zz += zK[i] * zK[i];
#pragma unroll
    for (int sh = 32; sh > 0; sh >>= 1) zz += __shfl_xor(zz, sh, 64);
    hi = rho * __shfl(zz, 0, 64);
    // guard against rounding in rho |z|^2
    while (secular_wave(dK, zK, k, dK[o], hi, rho, lane) < 0.0) hi *= 2.0;
  }
  const double dorg = dK[o];
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    if (secular_wave(dK, zK, k, dorg, mid, rho, lane) < 0.0) lo = mid;
    else hi = mid;
  }
  if (lane == 0) {
    org[off + jr] = o;
    tau_out[off + jr] = 0.5 * (lo + hi);
  }
}

// Gu-Eisenstat weights: zhat_i^2 = prod_j (lambda_j - d_i) / (rho prod_{j != i} (d_j - d_i)).
__global__ __launch_bounds__(kT) void dc_zhat_kernel(const Merge* merges, const int* cnt, const double* rho_in,
                                                     const double* dK_all, const double* zK_all, const int* org_all,
                                                     const double* tau_all, double* zh_all) {
  const Merge mg = merges[blockIdx.y];
  const int off = mg.off;
  const int k = cnt[3 * blockIdx.y];
  const int i = blockIdx.x * kT + threadIdx.x;
  if (i >= k) return;
  const double* dK = dK_all + off;
  const int* org = org_all + off;
  const double* tau = tau_all + off;
  const double di = dK[i];
  // lambda_j - d_i = (d_org(j) - d_i) + tau_j
  double p = ((dK[org[k - 1]] - di) + tau[k - 1]) / rho_in[blockIdx.y];
  for (int j = 0; j < i; ++j) p *= ((dK[org[j]] - di) + tau[j]) / (dK[j] - di);
  for (int j = i; j < k - 1; ++j) p *= ((dK[org[j]] - di) + tau[j]) / (dK[j + 1] - di);
  const double z = sqrt(fmax(p, 0.0));
  zh_all[off + i] = zK_all[off + i] >= 0.0 ? z : -z;
}

// Column norms of the secular eigenvectors u_j[i] = zhat_i / (d_i - lambda_j) (thread per root).
__global__ __launch_bounds__(kT) void dc_vnorm_kernel(const Merge* merges, const int* cnt, const double* dK_all,
                                                      const int* org_all, const double* tau_all,
                                                      const double* zh_all, double* inv_norm) {
  const Merge mg = merges[blockIdx.y];
  const int off = mg.off;
  const int k = cnt[3 * blockIdx.y];
  const int j = blockIdx.x * kT + threadIdx.x;
  if (j >= k) return;
  const double* dK = dK_all + off;
  const double dorg = dK[org_all[off + j]], tj = tau_all[off + j];
  double s = 0.0;
  for (int i = 0; i < k; ++i) {
    const double u = zh_all[off + i] / ((dK[i] - dorg) - tj);
    s += u * u;
  }
  inv_norm[off + j] = 1.0 / sqrt(s);
}

// New eigenvalues of the merged block (unsorted: roots then deflated), with
// the source of each output column, then ranks for the ascending order.
__global__ __launch_bounds__(kT) void dc_values_kernel(const Merge* merges, const int* cnt, const double* dK_all,
                                                       const int* org_all, const double* tau_all,
                                                       const double* dD_all, double* vals) {
  const Merge mg = merges[blockIdx.y];
  const int off = mg.off, s = mg.s1 + mg.s2;
  const int k = cnt[3 * blockIdx.y];
  const int q = blockIdx.x * kT + threadIdx.x;
  if (q >= s) return;
  double v;
  if (q < k) v = dK_all[off + org_all[off + q]] + tau_all[off + q];
  else v = dD_all[off + q - k];
  vals[off + q] = v;
}

// Output column rank[q] of value q (ascending; ties by index) and new lambda.
__global__ __launch_bounds__(kT) void dc_rank_kernel(const Merge* merges, const double* vals, int* rank,
                                                     double* lam_new) {
  const Merge mg = merges[blockIdx.y];
  const int off = mg.off, s = mg.s1 + mg.s2;
  const int q = blockIdx.x * kT + threadIdx.x;
  if (q >= s) return;
  const double x = vals[off + q];
  int r = 0;
  for (int p = 0; p < s; ++p) {
    const double y = vals[off + p];
    r += (y < x || (y == x && p < q)) ? 1 : 0;
  }
  rank[off + q] = r;
  lam_new[off + r] = x;
}

// U (block at off, ld ldu, zero outside the entries written here): output column rank[q]:
//   q <  k: sum_i u_q[i] e_{pK[i]}   (secular eigenvector)
//   q >= k: e_{pD[q-k]}              (deflated)
__global__ __launch_bounds__(kT) void dc_build_u_kernel(const Merge* merges, const int* cnt, const double* dK_all,
                                                        const int* pK_all, const int* pD_all, const int* org_all,
                                                        const double* tau_all, const double* zh_all,
                                                        const double* inv_norm, const int* rank, double* U,
                                                        int64_t ldu) {
  const Merge mg = merges[blockIdx.z];
  const int off = mg.off, s = mg.s1 + mg.s2;
  const int k = cnt[3 * blockIdx.z];
  const int q = blockIdx.y;  // source column
  if (q >= s) return;
  double* Ucol = U + off + (int64_t)(off + rank[off + q]) * ldu;  // zero-filled beforehand
  if (q < k) {
    const double* dK = dK_all + off;
    const double dorg = dK[org_all[off + q]], tq = tau_all[off + q], nrm = inv_norm[off + q];
    for (int i = blockIdx.x * kT + threadIdx.x; i < k; i += gridDim.x * kT)
      Ucol[pK_all[off + i]] = nrm * zh_all[off + i] / ((dK[i] - dorg) - tq);
  } else if (blockIdx.x == 0 && threadIdx.x == 0) {
    Ucol[pD_all[off + q - k]] = 1.0;
  }
}

// Leaves (size-1 blocks): Q = I on the diagonal, lambda = d - |e_left| - |e_right|.
__global__ __launch_bounds__(kT) void dc_leaves_kernel(const double* d, const double* e, int64_t n, double* lam,
                                                       double* Q, int64_t ldq) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  double x = d[i];
  if (i > 0) x -= fabs(e[i - 1]);
  if (i + 1 < n) x -= fabs(e[i]);
  lam[i] = x;
  Q[i + i * ldq] = 1.0;
}

// ---------------------------------------------------------------------------
// 3. Back-transformation helpers.
// ---------------------------------------------------------------------------
// V (mrows x jb, ld ldv): column c = reflector of A column k0 + c, rows from r0:
// unit at row k0 + c + 1, entries A[row][k0 + c] below, zero above.
__global__ __launch_bounds__(kT) void trd_extract_v_kernel(const double* A, int64_t ld, int64_t k0, int64_t r0,
                                                           int jb, int64_t mrows, double* V, int64_t ldv, int off) {
  const int64_t total = mrows * jb;
  for (int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x; e < total; e += (int64_t)gridDim.x * kT) {
    const int64_t i = e % mrows;
    const int c = (int)(e / mrows);
    const int64_t row = r0 + i, head = k0 + c + off;
    double x = 0.0;
    if (row == head) x = 1.0;
    else if (row > head) x = A[row + (k0 + c) * ld];
    V[i + (int64_t)c * ldv] = x;
  }
}

__global__ __launch_bounds__(kT) void fill2d_kernel(double* A, int64_t ld, int64_t m, int64_t n, double val) {
  for (int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x; e < m * n; e += (int64_t)gridDim.x * kT)
    A[(e % m) + (e / m) * ld] = val;
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
int ke_fill(double* A, int64_t ld, int64_t m, int64_t n, double val, hipStream_t s) {
  hipLaunchKernelGGL(fill2d_kernel, dim3((unsigned)std::min<int64_t>(8192, blocks_for(m * n))), dim3(kT), 0, s, A, ld,
                     m, n, val);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

int ke_trd_column(double* A, int64_t ld, int64_t n, int64_t k0, int64_t i, double* W, double* v, double* tau, double* e,
                  TrdScratch& w, bool last_in_panel, hipStream_t s) {
  const int64_t m = n - i;
  const int nwg = (int)std::min<int64_t>(64, std::max<int64_t>(1, (m + kT - 1) / kT));
  // the previous column of the panel left its W correction to this launch
  const int fix = (i > k0) ? 1 : 0;
  const int nwg_prev = (int)std::min<int64_t>(64, std::max<int64_t>(1, (m + 1 + kT - 1) / kT));
  hipLaunchKernelGGL(trd_col_kernel, dim3(nwg), dim3(kT), 0, s, A, ld, n, k0, i, W, w.part, w.sc, (const double*)v,
                     (const double*)w.wpart, nwg_prev, (const double*)tau, fix);
  if (i + 1 >= n) {
    ZOLO_HIP(hipGetLastError());
    return kOk;
  }
  hipLaunchKernelGGL(trd_house_kernel, dim3(nwg), dim3(kT), 0, s, A, ld, n, k0, i, W, w.part, nwg, w.sc, v, tau, e,
                     w.dpart);
  const int64_t m1 = n - i - 1;
  const int64_t T = (m1 + kSymT - 1) / kSymT;
  hipLaunchKernelGGL(trd_symv_kernel, dim3((unsigned)(T * (T + 1) / 2)), dim3(kT), 0, s, A, ld, n, i, v, w.ypart);
  hipLaunchKernelGGL(trd_w_kernel, dim3(nwg), dim3(kT), 0, s, A, ld, n, k0, i, W, v, w.ypart, (int)T, w.dpart, nwg, tau,
                     w.wpart);
  // the last column of a panel corrects its W column itself (the rank-2k
  // update follows); the others leave it to the next column's trd_col
  if (last_in_panel)
    hipLaunchKernelGGL(trd_wfix_kernel, dim3(nwg), dim3(kT), 0, s, W, n, (int)(i - k0), v, w.wpart, nwg, tau, i);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

int ke_trd_panel_end(double* A, int64_t ld, int64_t k0, int jb, int64_t n, double* d, hipStream_t s) {
  hipLaunchKernelGGL(trd_panel_end_kernel, dim3(1), dim3(64), 0, s, A, ld, k0, jb, n, d);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

int ke_trd_extract_v(const double* A, int64_t ld, int64_t k0, int64_t r0, int jb, int64_t mrows, double* V,
                     int64_t ldv, int off, hipStream_t s) {
  hipLaunchKernelGGL(trd_extract_v_kernel, dim3((unsigned)std::min<int64_t>(8192, blocks_for(mrows * jb))), dim3(kT), 0,
                     s, A, ld, k0, r0, jb, mrows, V, ldv, off);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

int ke_dc_leaves(const double* d, const double* e, int64_t n, double* lam, double* Q, int64_t ldq, hipStream_t s) {
  hipLaunchKernelGGL(dc_leaves_kernel, dim3(blocks_for(n)), dim3(kT), 0, s, d, e, n, lam, Q, ldq);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

int ke_dc_level(const DcLevel& L, hipStream_t s) {
  const int M = L.nmerge;
  const Merge* merges = reinterpret_cast<const Merge*>(L.merges);
  const unsigned gs = (unsigned)((L.smax + kT - 1) / kT);
  const unsigned tp = (unsigned)std::min(1024, ((L.smax + 63) / 64) * 64);
  hipLaunchKernelGGL(dc_prepare_kernel, dim3(M), dim3(tp), 0, s, merges, L.e, L.lam, L.Q, L.ldq, L.rho, L.dK, L.zK,
                     L.pK, L.dD, L.pD, L.rotcs, L.rotab, L.cnt, L.zbuf, L.dbuf, L.perm);
  hipLaunchKernelGGL(dc_rotate_kernel, dim3(gs, M), dim3(kT), 0, s, merges, L.Q, L.ldq, L.cnt, L.rotcs, L.rotab);
  const unsigned gr = (unsigned)((L.smax + kRootsPerBlock - 1) / kRootsPerBlock);
  hipLaunchKernelGGL(dc_roots_kernel, dim3(gr, M), dim3(kT), 0, s, merges, L.cnt, L.rho, L.dK, L.zK, L.org, L.tau);
  hipLaunchKernelGGL(dc_zhat_kernel, dim3(gs, M), dim3(kT), 0, s, merges, L.cnt, L.rho, L.dK, L.zK, L.org, L.tau,
                     L.zh);
  hipLaunchKernelGGL(dc_vnorm_kernel, dim3(gs, M), dim3(kT), 0, s, merges, L.cnt, L.dK, L.org, L.tau, L.zh,
                     L.inv_norm);
  hipLaunchKernelGGL(dc_values_kernel, dim3(gs, M), dim3(kT), 0, s, merges, L.cnt, L.dK, L.org, L.tau, L.dD,
                     L.vals);
  hipLaunchKernelGGL(dc_rank_kernel, dim3(gs, M), dim3(kT), 0, s, merges, L.vals, L.rank, L.lam_new);
  hipLaunchKernelGGL(dc_build_u_kernel, dim3(std::max(1u, gs), (unsigned)L.smax, M), dim3(kT), 0, s, merges, L.cnt,
                     L.dK, L.pK, L.pD, L.org, L.tau, L.zh, L.inv_norm, L.rank, L.U, L.ldu);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

}  // namespace zolo
