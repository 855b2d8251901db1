// Second stage of the two-stage tridiagonalisation that replaces ELPA's
// solve_evp_real_2stage (elpa_pdsyevd2.F90:199-200, called from
// ZoloCHK1.f90:86): the symmetric band matrix of bandwidth b left by the
// full -> band stage is reduced to tridiagonal form by bulge chasing, and the
// reflectors it generates are grouped for a blocked back-transformation.
//
// Band store: LB[c * ldb + d] = A(c + d, c), d < ldb = 2b (the lower band and
// room for the bulge, which reaches d = 2b - 1).
//
// Bulge chasing: sweep i (i = 0 .. n-3) clears column i below its
// subdiagonal.  Its task k works on the row block R_k = [s, s + L) with
// s = i + 1 + k b: the reflector H = I - tau v v^T is taken from column i
// (k = 0) or from the first column of the block E_{k-1} = A(R_k, R_{k-1}) that
// task k-1 filled in; H is applied from the left to the rest of E_{k-1}, on
// both sides of the diagonal block D_k = A(R_k, R_k), and from the right to
// E_k = A(R_{k+1}, R_k), which fills it in for task k+1.  Task (i+1, k)
// overlaps task (i, k+1) and shares one entry with task (i, k+2) (the head
// of the column task k+2 clears) and nothing with later tasks of sweep i, so
// the sweeps form a pipeline (see the kernel for the lag).  One workgroup
// runs one sweep at a time (sweeps i, i + W, i + 2W, ... for workgroup i of
// W), keeps E_k in LDS from task k to task k+1, and publishes its progress
// per sweep.
#include <algorithm>

#include "eig.hpp"

namespace zolo {

namespace {

constexpr int kT = 256;
constexpr int kLd = kSbNb + 1;  // LDS leading dimension of the 64 x 64 blocks

// The band store is shared by the workgroups of sb2st_kernel (each task reads
// what the previous sweep's workgroup wrote): inside that kernel every access
// to it is a device-scope load or store, which gfx950 performs at the
// device's coherence point (stores written through, loads not served from an
// incoherent cache).  A workgroup publishes a task after its stores have
// completed (vmcnt(0) in every thread, then a barrier), so neither side needs
// the L2 writeback and invalidate of a release/acquire pair.
__device__ inline double band_ld(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void band_st(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// s_waitcnt operand (gfx9 encoding): vmcnt(0), expcnt and lgkmcnt not waited for
constexpr int kWaitVm0 = 0x0F70;

__device__ inline double wsum(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}

__global__ __launch_bounds__(kT) void band_extract_kernel(const double* H, int64_t ldh, int64_t n, int b, double* LB,
                                                          int64_t ldb) {
  const int64_t total = n * ldb;
  for (int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x; e < total; e += (int64_t)gridDim.x * kT) {
    const int64_t d = e % ldb, c = e / ldb;
    LB[e] = (d <= b && c + d < n) ? H[(c + d) + c * ldh] : 0.0;
  }
}

__global__ __launch_bounds__(kT) void band_diag_kernel(const double* LB, int64_t ldb, int64_t n, double* d,
                                                       double* e) {
  const int64_t c = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (c >= n) return;
  d[c] = LB[c * ldb];
  e[c] = c + 1 < n ? LB[c * ldb + 1] : 0.0;
}

struct SbArgs {
  double* LB;
  int64_t ldb;
  int64_t n;
  int nsweep;
  int g;             // sweeps per group of the reflector store
  int kmax;          // tasks of sweep 0
  int64_t span;      // rows of a group block reflector: b + g - 1
  double* vstore;    // group (j, k): span x g at ((j * kmax + k) * g) * span
  double* tstore;    // tau of sweep j g + q, task k: (j * kmax + k) * g + q
  int* progress;     // tasks finished per sweep (zero on entry)
  int* status;       // 0, or 1 when a wait ran out of its spin budget
  long long spin_limit;
};

// B = the bandwidth (kSbNb).  Blocks in LDS are column-major with leading
// dimension kLd: X[u * kLd + t] = row t, column u of the block.  Thread
// (wave w, lane t) owns row t of the block columns u = w + 4 r, r < B / 4: it
// loads them (all loads of a task in flight at once), updates them in
// registers and stores them back.  The three products a task needs (D v,
// E_k v and v^T E_{k-1}) are partial sums over a quarter of the block per
// wave, combined through LDS.
//
// The reflector of task k+1 depends only on the first column of E_k, which is
// final once task k has applied H from the right (no other sweep touches it
// in between), so task k computes it at its end, stores the cleared column
// (beta e1) and only then publishes its progress.  The one entry that task
// (i+1, k) shares with task (i, k+2) is that column's head, so sweep i+1 may
// run task k once sweep i has published k + 2 (two tasks of lag instead of
// three).
template <int B>
__global__ __launch_bounds__(kT) void sb2st_kernel(SbArgs a) {
  static_assert(B == 64, "one lane per block row");
  constexpr int R = B / 4;            // block columns per thread
  __shared__ double blk[3][B * kLd];  // two E blocks (carried and new) and D (both triangles)
  __shared__ double vs[2][B], ws[B], qs[B], es[B], part[3][4][B];
  __shared__ double misc[2][2];       // tau, beta of the current and the next reflector
  __shared__ int flag;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t n = a.n, ldb = a.ldb;
  double* LB = a.LB;
  for (int i = blockIdx.x; i < a.nsweep; i += gridDim.x) {
    const int K = (int)((n - 2 - i) / B) + 1;
    const int Kprev = i > 0 ? (int)((n - 2 - (i - 1)) / B) + 1 : 0;
    double* Ep = blk[0];  // E_{k-1}, carried from the previous task
    double* En = blk[1];  // E_k
    double* Dm = blk[2];
    const int j = i / a.g, q = i % a.g;
    for (int k = 0; k < K; ++k) {
      const int cur = k & 1;
      const int64_t s = i + 1 + (int64_t)k * B;
      const int L = (int)std::min<int64_t>(B, n - s);
      const int64_t s2 = s + B;
      const int L2 = s2 < n ? (int)std::min<int64_t>(B, n - s2) : 0;
      // ---- wait until sweep i-1 has published task k+1 (finished, with the
      // cleared column of its task k+2)
      if (threadIdx.x == 0) {
        int ok = 1;
        if (i > 0) {
          const int need = std::min(k + 2, Kprev);
          long long spins = 0;
          while (__hip_atomic_load(a.progress + (i - 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
            if (__hip_atomic_load(a.status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
                ++spins > a.spin_limit) {
              __hip_atomic_store(a.status, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              ok = 0;
              break;
            }
            __builtin_amdgcn_s_sleep(1);
          }
        }
        flag = ok;
      }
      __syncthreads();  // (1)
      if (!flag) return;
      // ---- loads: D_k (lower triangle) and E_k in registers; for k = 0 the
      // column of sweep i and its reflector
      double dreg[R], ereg[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int u = wv + 4 * r;
        const bool ind = u < L && lane >= u && lane < L;
        const bool ine = u < L && lane < L2;
        dreg[r] = ind ? band_ld(LB + (s + u) * ldb + (lane - u)) : 0.0;
        ereg[r] = ine ? band_ld(LB + (s + u) * ldb + (B + lane - u)) : 0.0;
      }
      double x = 0.0;
      if (k == 0 && wv == 0 && lane < L) x = band_ld(LB + (int64_t)i * ldb + 1 + lane);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int u = wv + 4 * r;
        if (u < L && lane >= u && lane < L) {
          Dm[u * kLd + lane] = dreg[r];
          Dm[lane * kLd + u] = dreg[r];
        }
        En[u * kLd + lane] = ereg[r];
      }
      if (k == 0 && wv == 0) {
        // H x = beta e1, LAPACK dlarfg without rescaling
        const double alpha = __shfl(x, 0, 64);
        const double sig = wsum(lane > 0 ? x * x : 0.0);
        double tau = 0.0, beta = alpha, v = lane == 0 ? 1.0 : 0.0;
        if (sig > 0.0) {
          beta = -copysign(sqrt(alpha * alpha + sig), alpha);
          tau = (beta - alpha) / beta;
          v = lane == 0 ? 1.0 : x * (1.0 / (alpha - beta));
        }
        vs[0][lane] = lane < L ? v : 0.0;
        if (lane < L) band_st(LB + (int64_t)i * ldb + 1 + lane, lane == 0 ? beta : 0.0);
        if (lane == 0) {
          misc[0][0] = tau;
          misc[0][1] = beta;
          a.tstore[(int64_t)j * a.kmax * a.g + q] = tau;
        }
        if (lane < L) a.vstore[((int64_t)j * a.kmax * a.g + q) * a.span + q + lane] = vs[0][lane];
      }
      __syncthreads();  // (2)
      const double tau = misc[cur][0], beta = misc[cur][1];
      const double vl = vs[cur][lane];
      double vu[R];  // v at this thread's block columns
#pragma unroll
      for (int r = 0; r < R; ++r) vu[r] = vs[cur][wv + 4 * r];
      // ---- partial sums of D v, E_k v (rows: lane; this wave's columns) and
      // v^T E_{k-1} (columns: lane; this wave's quarter of the rows).  Every
      // LDS read is issued before the first product, so the reads overlap
      // instead of each waiting for the one before.  Entries outside the
      // task (rows or columns >= L, E_{k-1} at k = 0) hold stale values; they
      // are dropped by selects, never multiplied by zero.
      {
        double dv[R], ev[R], vt[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          dv[r] = Dm[(wv + 4 * r) * kLd + lane];
          ev[r] = Ep[lane * kLd + wv * R + r];
          vt[r] = vs[cur][wv * R + r];
        }
        double p = 0.0, rr = 0.0, ee = 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          p += (wv + 4 * r < L && lane < L) ? dv[r] * vu[r] : 0.0;
          rr += ereg[r] * vu[r];  // ereg and vu are zero outside the task
          ee += vt[r] * ev[r];
        }
        part[0][wv][lane] = p;
        part[1][wv][lane] = rr;
        part[2][wv][lane] = k > 0 ? ee : 0.0;
      }
      __syncthreads();  // (3)
      if (wv == 0) {
        const double p = lane < L ? tau * (part[0][0][lane] + part[0][1][lane] + part[0][2][lane] + part[0][3][lane])
                                  : 0.0;
        const double pv = wsum(p * vl);
        ws[lane] = p - 0.5 * tau * pv * vl;
      } else if (wv == 1) {
        qs[lane] = lane < L2 ? tau * (part[1][0][lane] + part[1][1][lane] + part[1][2][lane] + part[1][3][lane])
                             : 0.0;
      } else if (wv == 2) {
        es[lane] = tau * (part[2][0][lane] + part[2][1][lane] + part[2][2][lane] + part[2][3][lane]);
      }
      __syncthreads();  // (4)
      // ---- D <- H D H (own lower entries, stored), E_{k-1} <- H E_{k-1}
      // (column 0 is beta e1) stored, E_k <- E_k H kept in LDS for task k+1;
      // wave 0 holds column 0 of E_k and takes the next reflector from it.
      // The LDS operands are read first, and the band stores of D and E_{k-1}
      // are issued before wave 0 forms the reflector, so they are in flight
      // while it computes.
      const double wl = ws[lane], ql = qs[lane];
      double wu[R], eu[R], pu[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int u = wv + 4 * r;
        wu[r] = ws[u];
        eu[r] = es[u];
        pu[r] = Ep[u * kLd + lane];
      }
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (wv + 4 * r < L && lane < L2) ereg[r] -= ql * vu[r];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int u = wv + 4 * r;
        if (u < L && lane >= u && lane < L) band_st(LB + (s + u) * ldb + (lane - u), dreg[r] - (vl * wu[r] + wl * vu[r]));
        if (k > 0 && lane < L) {
          const double y = u == 0 ? (lane == 0 ? beta : 0.0) : pu[r] - vl * eu[r];
          band_st(LB + (s - B + u) * ldb + (B + lane - u), y);
        }
      }
      if (wv == 0 && L2 > 0) {
        const int nx = cur ^ 1;
        const double xx = lane < L2 ? ereg[0] : 0.0;
        const double alpha = __shfl(xx, 0, 64);
        const double sig = wsum(lane > 0 ? xx * xx : 0.0);
        double tn = 0.0, bn = alpha, v = lane == 0 ? 1.0 : 0.0;
        if (sig > 0.0) {
          bn = -copysign(sqrt(alpha * alpha + sig), alpha);
          tn = (bn - alpha) / bn;
          v = lane == 0 ? 1.0 : xx * (1.0 / (alpha - bn));
        }
        vs[nx][lane] = lane < L2 ? v : 0.0;
        ereg[0] = lane == 0 ? bn : 0.0;
        if (lane < L2) band_st(LB + s * ldb + B + lane, ereg[0]);  // column 0 of E_k, cleared
        const int64_t grp = (int64_t)j * a.kmax + (k + 1);
        if (lane < L2) a.vstore[(grp * a.g + q) * a.span + q + lane] = v;
        if (lane == 0) {
          misc[nx][0] = tn;
          misc[nx][1] = bn;
          a.tstore[grp * a.g + q] = tn;
        }
      }
#pragma unroll
      for (int r = 0; r < R; ++r) En[(wv + 4 * r) * kLd + lane] = ereg[r];
      // this thread's band stores are complete before the barrier, so all of
      // the task's are before the progress store
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      __builtin_amdgcn_s_waitcnt(kWaitVm0);
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      __syncthreads();  // (5)
      if (threadIdx.x == 0)
        __hip_atomic_store(a.progress + i, k + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      double* tmp = Ep;
      Ep = En;
      En = tmp;
    }
  }
}

// T of every group block reflector (j, k) from G = V^T V (computed by the
// GEMM engine): V is span x g with column q the reflector of sweep j g + q at
// rows q .., and H_0 H_1 ... H_{g-1} = I - V T V^T with T upper triangular
// (LAPACK dlarft, forward, columnwise): T(q, q) = tau_q and T(0:q, q) =
// -tau_q T(0:q, 0:q) G(0:q, q).  One workgroup per group, thread p holds row p.
template <int G>
__global__ __launch_bounds__(G) void q2_form_t_kernel(const double* Gstore, const double* tstore, int ngroups,
                                                      double* Tstore) {
  __shared__ double Tl[G * (G + 1)];
  const int grp = blockIdx.x;
  if (grp >= ngroups) return;
  const double* Gm = Gstore + (int64_t)grp * G * G;
  const double* tau = tstore + (int64_t)grp * G;
  double* Tm = Tstore + (int64_t)grp * G * G;
  const int p = threadIdx.x;
  for (int qq = 0; qq < G; ++qq) {
    const double tq = tau[qq];
    double y = 0.0;
    for (int c = p; c < qq; ++c) y += Tl[c * (G + 1) + p] * Gm[qq * G + c];
    Tl[qq * (G + 1) + p] = p < qq ? -tq * y : (p == qq ? tq : 0.0);
    __syncthreads();
  }
  for (int c = 0; c < G; ++c) Tm[c * G + p] = Tl[c * (G + 1) + p];
}

}  // namespace

int ke_band_extract(const double* H, int64_t ldh, int64_t n, int b, double* LB, int64_t ldb, hipStream_t s) {
  const int64_t total = n * ldb;
  const unsigned g = (unsigned)std::min<int64_t>(8192, (total + kT - 1) / kT);
  hipLaunchKernelGGL(band_extract_kernel, dim3(g), dim3(kT), 0, s, H, ldh, n, b, LB, ldb);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

int ke_band_diag(const double* LB, int64_t ldb, int64_t n, double* d, double* e, hipStream_t s) {
  hipLaunchKernelGGL(band_diag_kernel, dim3((unsigned)((n + kT - 1) / kT)), dim3(kT), 0, s, LB, ldb, n, d, e);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

int ke_sb2st(double* LB, int64_t ldb, int64_t n, int g, int kmax, double* vstore, double* tstore, int* progress,
             int* status, int nwg, hipStream_t s) {
  if (n < 3) return kOk;
  if (ldb < 2 * kSbNb || g != kSbGroup) {
    set_error("sb2st: band store or group size does not match the kernel");
    return kErrArg;
  }
  SbArgs a;
  a.LB = LB;
  a.ldb = ldb;
  a.n = n;
  a.nsweep = (int)(n - 2);
  a.g = g;
  a.kmax = kmax;
  a.span = kSbNb + g - 1;
  a.vstore = vstore;
  a.tstore = tstore;
  a.progress = progress;
  a.status = status;
  a.spin_limit = 1ll << 22;  // a few seconds of polling at most
  const int w = std::max(1, std::min(nwg, a.nsweep));
  // Every workgroup waits on the one running the previous sweep, so all of
  // them must be resident at once: a cooperative launch guarantees it.
  void* args[] = {&a};
  ZOLO_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&sb2st_kernel<kSbNb>), dim3(w), dim3(kT), args, 0,
                                      s));
  return kOk;
}

int ke_q2_form_t(const double* Gstore, const double* tstore, int ngroups, double* Tstore, hipStream_t s) {
  if (ngroups <= 0) return kOk;
  hipLaunchKernelGGL((q2_form_t_kernel<kSbGroup>), dim3(ngroups), dim3(kSbGroup), 0, s, Gstore, tstore, ngroups,
                     Tstore);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

}  // namespace zolo
