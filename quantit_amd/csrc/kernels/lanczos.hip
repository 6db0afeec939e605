// Kernels of the k-step Lanczos eigensolver of the two-site update (FP64).
//
// The Krylov vectors of one update share one flat layout (the blocks of the
// two-site state, back to back), so they are the rows of one matrix V
// (row pitch ld): the full reorthogonalisation is a pair of skinny
// products c = V w and w -= V^T c, and the Ritz vector is V^T y.  All
// reductions run in a fixed order (no atomics), so results do not depend on
// what else runs on the device.
#include "../kernels.hpp"

namespace qtt {
namespace kern {

namespace {

constexpr int kThreads = 256;

__device__ double block_reduce(double v, double* sh) {
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (tid < w) sh[tid] += sh[tid + w];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

}  // namespace

// partial[i * G + b] = sum over chunk b of V[i][e] * w[e]; the thread keeps
// its kKrylovPerThread elements of w in registers across the m rows.
__global__ __launch_bounds__(kThreads) void k_krylov_dot_partial(double* __restrict__ base, int64_t v_off, int64_t ld,
                                                                 int m, int64_t w_off, int64_t n,
                                                                 int64_t partial_off) {
  __shared__ double sh[kThreads];
  const int G = gridDim.x;
  const int64_t e0 = int64_t(blockIdx.x) * kKrylovChunk + threadIdx.x;
  const double* w = base + w_off;
  double wr[kKrylovPerThread];
#pragma unroll
  for (int q = 0; q < kKrylovPerThread; ++q) {
    const int64_t e = e0 + int64_t(q) * kThreads;
    wr[q] = e < n ? w[e] : 0.0;
  }
  for (int i = 0; i < m; ++i) {
    const double* v = base + v_off + int64_t(i) * ld;
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < kKrylovPerThread; ++q) {
      const int64_t e = e0 + int64_t(q) * kThreads;
      if (e < n) acc += v[e] * wr[q];
    }
    const double r = block_reduce(acc, sh);
    if (threadIdx.x == 0) base[partial_off + int64_t(i) * G + blockIdx.x] = r;
  }
}

// out[i] = sum_b partial[i * G + b], b ascending.
__global__ void k_krylov_dot_final(double* __restrict__ base, int64_t partial_off, int G, int m, int64_t out_off) {
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    double acc = 0.0;
    const double* p = base + partial_off + int64_t(i) * G;
    for (int b = 0; b < G; ++b) acc += p[b];
    base[out_off + i] = acc;
  }
}

void krylov_dot(hipStream_t s, double* base, int64_t v_off, int64_t ld, int m, int64_t w_off, int64_t n,
                int64_t partial_off, int64_t out_off) {
  if (m <= 0) return;
  const int G = int((n + kKrylovChunk - 1) / kKrylovChunk);
  if (G > 0)
    hipLaunchKernelGGL(k_krylov_dot_partial, dim3(G), dim3(kThreads), 0, s, base, v_off, ld, m, w_off, n,
                       partial_off);
  hipLaunchKernelGGL(k_krylov_dot_final, dim3(1), dim3(64), 0, s, base, partial_off, G, m, out_off);
}

// dst[e] = src[e] + sign * sum_i coef[i] V[i][e] (src_off < 0: no src term;
// scale_off >= 0: the result is divided by sqrt(base[scale_off])).
__global__ __launch_bounds__(kThreads) void k_krylov_axpy(double* __restrict__ base, int64_t v_off, int64_t ld, int m,
                                                          int64_t coef_off, double sign, int64_t src_off,
                                                          int64_t dst_off, int64_t n, int64_t scale_off) {
  const int64_t e = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (e >= n) return;
  double acc = 0.0;
  for (int i = 0; i < m; ++i) acc += base[coef_off + i] * base[v_off + int64_t(i) * ld + e];
  double r = (src_off >= 0 ? base[src_off + e] : 0.0) + sign * acc;
  if (scale_off >= 0) r = r / sqrt(base[scale_off]);
  base[dst_off + e] = r;
}

void krylov_axpy(hipStream_t s, double* base, int64_t v_off, int64_t ld, int m, int64_t coef_off, double sign,
                 int64_t src_off, int64_t dst_off, int64_t n, int64_t scale_off) {
  if (n <= 0) return;
  const int64_t G = (n + kThreads - 1) / kThreads;
  hipLaunchKernelGGL(k_krylov_axpy, dim3(unsigned(G)), dim3(kThreads), 0, s, base, v_off, ld, m, coef_off, sign,
                     src_off, dst_off, n, scale_off);
}

namespace {

// Number of eigenvalues of the symmetric tridiagonal T (alpha, beta) below
// x: the negative pivots of the LDL^T factorisation of T - x I (Sturm).
__device__ int sturm_count(const double* alpha, const double* beta, int m, double x, double tiny) {
  int cnt = 0;
  double q = alpha[0] - x;
  if (q == 0.0) q = -tiny;
  cnt += q < 0.0;
  for (int i = 1; i < m; ++i) {
    q = alpha[i] - x - beta[i - 1] * beta[i - 1] / q;
    if (q == 0.0) q = -tiny;
    cnt += q < 0.0;
  }
  return cnt;
}

}  // namespace

// One wave.  Stores alpha[j] = c1[j] + c2[j] and beta[j] = sqrt(nrm2) into
// T (alpha at T[0..K), beta at T[K..2K)), then the lowest eigenpair of the
// leading (j+1) x (j+1) tridiagonal:
// * eigenvalue by multisection: 64 lanes evaluate Sturm counts at 64 interior
//   points of the bracket, which shrinks 65-fold per round, from the
//   Gershgorin interval down to a few ulps;
// * eigenvector by two inverse iterations with T - sigma I, sigma just below
//   the eigenvalue (positive definite: its LDL^T pivots stay positive),
//   normalised with y[0] >= 0.
// out: [0] eigenvalue, [1] residual norm beta[j] |y[j]|, [2] beta[j],
// [3 .. 3 + j] y.  flags[0] = 1 when a NaN appears.
__global__ __launch_bounds__(64) void k_lanczos_tridiag(double* __restrict__ base, int64_t t_off, int K, int j,
                                                        int64_t c1_off, int64_t c2_off, int64_t nrm2_off,
                                                        int64_t out_off, int* flags) {
  __shared__ double alpha[kMaxKrylov], beta[kMaxKrylov], y[kMaxKrylov], l[kMaxKrylov], dd[kMaxKrylov];
  __shared__ double bracket[2];
  __shared__ int first[1];
  const int lane = threadIdx.x;
  const int m = j + 1;
  if (lane == 0) {
    base[t_off + j] = base[c1_off + j] + base[c2_off + j];
    base[t_off + K + j] = sqrt(base[nrm2_off]);
  }
  __syncthreads();
  for (int i = lane; i < m; i += 64) {
    alpha[i] = base[t_off + i];
    beta[i] = base[t_off + K + i];
  }
  __syncthreads();
  if (lane == 0) {
    double lo = alpha[0], hi = alpha[0], scale = 0.0;
    for (int i = 0; i < m; ++i) {
      const double r = (i > 0 ? fabs(beta[i - 1]) : 0.0) + (i + 1 < m ? fabs(beta[i]) : 0.0);
      lo = fmin(lo, alpha[i] - r);
      hi = fmax(hi, alpha[i] + r);
      scale = fmax(scale, fabs(alpha[i]) + r);
    }
    const double pad = 2.2e-16 * (scale + 1e-300);
    bracket[0] = lo - pad;
    bracket[1] = hi + pad;
  }
  __syncthreads();
  double scale = fmax(fabs(bracket[0]), fabs(bracket[1]));
  const double tiny = 1e-300 + 2.2e-16 * 1e-30 * scale;
  // invariant: count(bracket[0]) == 0, count(bracket[1]) >= 1
  for (int round = 0; round < 12; ++round) {
    const double lo = bracket[0], hi = bracket[1];
    if (!(hi - lo > 4.4e-16 * fmax(fabs(lo), fabs(hi)))) break;
    const double x = lo + (hi - lo) * double(lane + 1) / 65.0;
    const int c = sturm_count(alpha, beta, m, x, tiny);
    const unsigned long long ge = __ballot(c >= 1);
    __syncthreads();
    if (lane == 0) {
      if (ge == 0ull) {
        bracket[0] = lo + (hi - lo) * 64.0 / 65.0;
      } else {
        const int f = __ffsll(static_cast<long long>(ge)) - 1;  // first lane whose point holds an eigenvalue below
        bracket[1] = lo + (hi - lo) * double(f + 1) / 65.0;
        if (f > 0) bracket[0] = lo + (hi - lo) * double(f) / 65.0;
      }
    }
    __syncthreads();
  }
  if (lane == 0) {
    const double lam = 0.5 * (bracket[0] + bracket[1]);
    // sigma below lam: T - sigma I positive definite (pivots > 0)
    const double sigma = lam - fmax(8.0 * (bracket[1] - bracket[0]), 4.4e-16 * fmax(scale, 1e-300));
    dd[0] = alpha[0] - sigma;
    for (int i = 1; i < m; ++i) {
      l[i - 1] = beta[i - 1] / dd[i - 1];
      dd[i] = alpha[i] - sigma - l[i - 1] * beta[i - 1];
      if (!(dd[i] > 0.0)) dd[i] = 1e-300 + 2.2e-16 * scale;
    }
    for (int i = 0; i < m; ++i) y[i] = 1.0;
    for (int it = 0; it < 3; ++it) {
      for (int i = 1; i < m; ++i) y[i] -= l[i - 1] * y[i - 1];   // L z = y
      for (int i = 0; i < m; ++i) y[i] /= dd[i];                  // D
      for (int i = m - 2; i >= 0; --i) y[i] -= l[i] * y[i + 1];  // L^T
      double nrm = 0.0;
      for (int i = 0; i < m; ++i) nrm += y[i] * y[i];
      nrm = sqrt(nrm);
      const double sgn = y[0] < 0.0 ? -1.0 : 1.0;
      for (int i = 0; i < m; ++i) y[i] = sgn * y[i] / nrm;
    }
    const double bj = beta[j];
    const double res = fabs(bj * y[m - 1]);
    base[out_off + 0] = lam;
    base[out_off + 1] = res;
    base[out_off + 2] = bj;
    bool bad = isnan(lam) || isnan(res);
    for (int i = 0; i < m; ++i) {
      base[out_off + 3 + i] = y[i];
      bad |= isnan(y[i]);
    }
    if (bad) flags[0] = 1;
  }
}

void lanczos_tridiag(hipStream_t s, double* base, int64_t t_off, int K, int j, int64_t c1_off, int64_t c2_off,
                     int64_t nrm2_off, int64_t out_off, int* flags) {
  hipLaunchKernelGGL(k_lanczos_tridiag, dim3(1), dim3(64), 0, s, base, t_off, K, j, c1_off, c2_off, nrm2_off,
                     out_off, flags);
}

}  // namespace kern
}  // namespace qtt
