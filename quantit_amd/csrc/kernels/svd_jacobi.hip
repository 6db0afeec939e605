// Batched FP64 block SVD for gfx950: block one-sided (Hestenes) Jacobi.
//
// Each sector matrix is held as Y (ell x kp, column-major) with
// ell = max(m,n) rows and kp >= min(m,n) columns (zero padded to a multiple of
// 2*kSvdBlock); W (kp x kp) accumulates the right rotations (Y W = Z with
// mutually orthogonal columns).  One launch = one round of the round-robin
// tournament over column blocks of kSvdBlock columns: every work-group takes
// a disjoint block pair (p, q) of one sector and
//   1. forms the 32x32 Gram matrix G = [Y_p Y_q]^T [Y_p Y_q] with
//      v_mfma_f64_16x16x4_f64 (Y staged through LDS in 64-row chunks),
//   2. returns if every column pair is orthogonal to the tolerance
//      (|G_ij| <= tol sqrt(G_ii G_jj)),
//   3. diagonalises G with a cyclic two-sided Jacobi in LDS (16 disjoint
//      rotations per step, circle ordering) accumulating Q,
//   4. applies Y_pq <- Y_pq Q and W_pq <- W_pq Q with MFMA.
// The singular values are the final column norms of Y (computed from the
// columns, so the small ones keep absolute accuracy eps*||A||), which matches
// the accuracy class of LAPACK gesdd used by the reference
// (sources/btensor_linalg.cpp:483).
#include <cfloat>

#include "../kernels.hpp"

namespace qtt {
namespace kern {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int NB2 = 2 * kSvdBlock;  // 32 columns per work-group
constexpr int YP = 66;              // LDS pitch of the staged Y chunk
constexpr int GP = NB2 + 1;         // LDS pitch of G and Q

__global__ void k_svd_init_w(double* __restrict__ base, const SvdSector* __restrict__ secs) {
  const SvdSector s = secs[blockIdx.x];
  const int kp = s.nb * kSvdBlock;
  double* W = base + s.w_off;
  const int64_t n = int64_t(kp) * kp;
  for (int64_t e = threadIdx.x; e < n; e += blockDim.x) {
    const int64_t r = e % kp, c = e / kp;
    W[e] = (r == c) ? 1.0 : 0.0;
  }
}

void svd_init_w(hipStream_t st, double* base, const SvdSector* secs, int nsec) {
  if (nsec <= 0) return;
  hipLaunchKernelGGL(k_svd_init_w, dim3(nsec), dim3(256), 0, st, base, secs);
}

__device__ __forceinline__ int circle_block(int pos, int round, int nb) {
  return pos == 0 ? 0 : 1 + ((pos - 1 + round) % (nb - 1));
}

// Apply the 32x32 rotation Q (LDS, pitch GP) to the 32 selected columns of a
// column-major matrix M (rows x ld) in place, 64 rows at a time.
__device__ void apply_rotation(double* __restrict__ M, int64_t ld, int rows, const int* cols,
                               double (*Ys)[YP], const double (*Q)[GP]) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int l0 = 0; l0 < rows; l0 += 64) {
    {
      const int c = tid >> 3, lq = (tid & 7) * 8;
      const double* src = M + int64_t(cols[c]) * ld;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int l = l0 + lq + i;
        Ys[c][lq + i] = (l < rows) ? src[l] : 0.0;
      }
    }
    __syncthreads();
    v4d acc[2] = {v4d{0.0, 0.0, 0.0, 0.0}, v4d{0.0, 0.0, 0.0, 0.0}};
#pragma unroll
    for (int kk = 0; kk < NB2; kk += 4) {
      const int kr = kk + (lane >> 4), c = lane & 15;
      const double a = Ys[kr][wave * 16 + c];
      const double b0 = Q[kr][c], b1 = Q[kr][16 + c];
      acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b0, acc[0], 0, 0, 0);
      acc[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b1, acc[1], 0, 0, 0);
    }
    __syncthreads();  // every wave has read its Ys rows before they are overwritten
#pragma unroll
    for (int jt = 0; jt < 2; ++jt)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = l0 + wave * 16 + 4 * r + (lane >> 4);  // f64 MFMA output layout
        const int col = jt * 16 + (lane & 15);
        if (row < rows) M[int64_t(cols[col]) * ld + row] = acc[jt][r];
      }
  }
}

__global__ __launch_bounds__(256) void k_svd_round(double* __restrict__ base, const SvdSector* __restrict__ secs,
                                                   const int32_t* __restrict__ work, int round,
                                                   int* __restrict__ counters, double tol) {
  __shared__ double Ys[NB2][YP];
  __shared__ double G[2][NB2][GP];
  __shared__ double Q[2][NB2][GP];
  __shared__ double cc[NB2], sg[NB2];
  __shared__ int partner[NB2];
  __shared__ int cols[NB2];
  __shared__ int flag, nrot;

  const SvdSector s = secs[work[2 * blockIdx.x]];
  const int pi = work[2 * blockIdx.x + 1];
  const int nb = s.nb;
  if (round >= nb - 1) return;
  const int p = circle_block(pi, round, nb);
  const int q = circle_block(nb - 1 - pi, round, nb);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid < NB2) cols[tid] = tid < kSvdBlock ? p * kSvdBlock + tid : q * kSvdBlock + (tid - kSvdBlock);
  if (tid == 0) flag = 0;
  __syncthreads();

  double* Y = base + s.y_off;
  const int64_t ld = s.ell;
  const int ell = s.ell;
  // ---- 1. Gram matrix ------------------------------------------------------
  const int wi = wave >> 1, wj = wave & 1;
  v4d g = v4d{0.0, 0.0, 0.0, 0.0};
  for (int l0 = 0; l0 < ell; l0 += 64) {
    {
      const int c = tid >> 3, lq = (tid & 7) * 8;
      const double* src = Y + int64_t(cols[c]) * ld;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int l = l0 + lq + i;
        Ys[c][lq + i] = (l < ell) ? src[l] : 0.0;
      }
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 64; kk += 4) {
      const int kr = kk + (lane >> 4), c = lane & 15;
      g = __builtin_amdgcn_mfma_f64_16x16x4f64(Ys[wi * 16 + c][kr], Ys[wj * 16 + c][kr], g, 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) G[0][wi * 16 + 4 * r + (lane >> 4)][wj * 16 + (lane & 15)] = g[r];
  __syncthreads();
  // ---- 2. convergence test ---------------------------------------------------
  for (int e = tid; e < NB2 * NB2; e += blockDim.x) {
    const int i = e / NB2, j = e % NB2;
    if (i < j) {
      const double gii = G[0][i][i], gjj = G[0][j][j], gij = G[0][i][j];
      if (gii > 0.0 && gjj > 0.0 && fabs(gij) > tol * sqrt(gii) * sqrt(gjj)) flag = 1;
    }
  }
  __syncthreads();
  if (flag == 0) return;
  if (tid == 0) atomicAdd(&counters[s.flag_idx], 1);
  // ---- 3. eigen-decomposition of G by cyclic Jacobi --------------------------
  for (int e = tid; e < NB2 * NB2; e += blockDim.x) {
    const int i = e / NB2, j = e % NB2;
    Q[0][i][j] = (i == j) ? 1.0 : 0.0;
  }
  int cur = 0;
  __syncthreads();
  for (int sweep = 0; sweep < 12; ++sweep) {
    if (tid == 0) nrot = 0;
    __syncthreads();
    for (int r = 0; r < NB2 - 1; ++r) {
      if (tid < NB2 / 2) {
        const int a = circle_block(tid, r, NB2), b = circle_block(NB2 - 1 - tid, r, NB2);
        const int pp = min(a, b), qq = max(a, b);
        const double app = G[cur][pp][pp], aqq = G[cur][qq][qq], apq = G[cur][pp][qq];
        double c = 1.0, sn = 0.0;
        if (app > 0.0 && aqq > 0.0 && fabs(apq) > DBL_EPSILON * sqrt(app) * sqrt(aqq)) {
          const double tau = (aqq - app) / (2.0 * apq);
          const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
          c = 1.0 / sqrt(1.0 + t * t);
          sn = c * t;
          atomicAdd(&nrot, 1);
        }
        partner[pp] = qq;
        partner[qq] = pp;
        cc[pp] = c;
        cc[qq] = c;
        sg[pp] = -sn;
        sg[qq] = sn;
      }
      __syncthreads();
      const int nx = cur ^ 1;
      for (int e = tid; e < NB2 * NB2; e += blockDim.x) {
        const int a = e / NB2, b = e % NB2;
        const int a2 = partner[a], b2 = partner[b];
        const double ca = cc[a], sa = sg[a], cb = cc[b], sb = sg[b];
        G[nx][a][b] = ca * cb * G[cur][a][b] + ca * sb * G[cur][a][b2] + sa * cb * G[cur][a2][b] +
                      sa * sb * G[cur][a2][b2];
        Q[nx][a][b] = cb * Q[cur][a][b] + sb * Q[cur][a][b2];
      }
      cur = nx;
      __syncthreads();
    }
    if (nrot == 0) break;
    __syncthreads();
  }
  // ---- 4. rotate the columns of Y and W ---------------------------------------
  apply_rotation(Y, ld, ell, cols, Ys, Q[cur]);
  __syncthreads();
  const int kp = nb * kSvdBlock;
  apply_rotation(base + s.w_off, kp, kp, cols, Ys, Q[cur]);
}

void svd_round(hipStream_t st, double* base, const SvdSector* secs, const int32_t* work, int nwork, int round,
               int* counters, double tol) {
  if (nwork <= 0) return;
  hipLaunchKernelGGL(k_svd_round, dim3(nwork), dim3(256), 0, st, base, secs, work, round, counters, tol);
}

// Block-wide reductions (fixed order, deterministic).
__device__ double wg_sum(double v, double* sh) {
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (tid < w) sh[tid] += sh[tid + w];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// One work-group per sector: singular values (column norms of Y), descending
// order, normalised Z columns and W columns scattered to U and V (row-major,
// m x k and n x k), and an orthonormal completion of the columns whose
// singular value is numerically zero.  aux = base + w_off + kp*kp holds
// 2*kp + ell doubles of scratch.
__global__ __launch_bounds__(256) void k_svd_finalize(double* __restrict__ base, const SvdSector* __restrict__ secs) {
  __shared__ double sh[256];
  const SvdSector s = secs[blockIdx.x];
  const int k = s.k, ell = s.ell, kp = s.nb * kSvdBlock;
  const double* Y = base + s.y_off;
  const double* W = base + s.w_off;
  double* aux = base + s.w_off + int64_t(kp) * kp;
  double* sig = aux;            // column norms
  double* rnk = aux + kp;       // rank of each column
  double* v = aux + 2 * kp;     // completion vector (ell)
  double* S = base + s.s_off;
  const int tid = threadIdx.x;
  // column norms: one column per thread group of 64 (wave)
  const int wave = tid >> 6, lane = tid & 63;
  for (int j = wave; j < k; j += 4) {
    double acc = 0.0;
    for (int l = lane; l < ell; l += 64) {
      const double x = Y[int64_t(j) * ell + l];
      acc += x * x;
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) sig[j] = sqrt(acc);
  }
  __syncthreads();
  for (int j = tid; j < k; j += blockDim.x) {
    const double sj = sig[j];
    int r = 0;
    for (int i = 0; i < k; ++i) {
      const double si = sig[i];
      r += (si > sj) || (si == sj && i < j);
    }
    rnk[j] = double(r);
    if (s.eig_shift)
      S[k - 1 - r] = sj - base[s.shift_off];
    else
      S[r] = sj;
  }
  __syncthreads();
  if (s.eig_shift) {
    // eigenvectors: U[row, asc] = W[row, j]
    double* U = base + s.u_off;
    const int64_t n = int64_t(s.m) * k;
    for (int64_t e = tid; e < n; e += blockDim.x) {
      const int row = int(e / k), j = int(e % k);
      const int asc = k - 1 - int(rnk[j]);
      U[int64_t(row) * k + asc] = W[int64_t(j) * kp + row];
    }
    return;
  }
  const double smax = k > 0 ? S[0] : 0.0;
  const double null_thr = smax * double(ell) * DBL_EPSILON;
  double* Zout = base + (s.transposed ? s.v_off : s.u_off);  // ell x k
  double* Wout = base + (s.transposed ? s.u_off : s.v_off);  // k x k (rows = min dim)
  {
    const int64_t n = int64_t(ell) * k;
    for (int64_t e = tid; e < n; e += blockDim.x) {
      const int row = int(e / k), j = int(e % k);
      const int r = int(rnk[j]);
      const double sj = sig[j];
      Zout[int64_t(row) * k + r] = (sj > null_thr && sj > 0.0) ? Y[int64_t(j) * ell + row] / sj : 0.0;
    }
    const int64_t nw = int64_t(k) * k;
    for (int64_t e = tid; e < nw; e += blockDim.x) {
      const int row = int(e / k), j = int(e % k);
      Wout[int64_t(row) * k + int(rnk[j])] = W[int64_t(j) * kp + row];
    }
  }
  __syncthreads();
  // orthonormal completion of null columns, in rank order
  int first_null = k;
  for (int t = 0; t < k; ++t)
    if (!(S[t] > null_thr && S[t] > 0.0)) {
      first_null = t;
      break;
    }
  for (int t = first_null; t < k; ++t) {
    for (int cand = 0; cand < ell; ++cand) {
      const int e0 = (t + cand) % ell;
      for (int l = tid; l < ell; l += blockDim.x) v[l] = (l == e0) ? 1.0 : 0.0;
      __syncthreads();
      for (int pass = 0; pass < 2; ++pass)
        for (int c = 0; c < t; ++c) {
          double acc = 0.0;
          for (int l = tid; l < ell; l += blockDim.x) acc += Zout[int64_t(l) * k + c] * v[l];
          const double pr = wg_sum(acc, sh);
          for (int l = tid; l < ell; l += blockDim.x) v[l] -= pr * Zout[int64_t(l) * k + c];
          __syncthreads();
        }
      double acc = 0.0;
      for (int l = tid; l < ell; l += blockDim.x) acc += v[l] * v[l];
      const double nrm = sqrt(wg_sum(acc, sh));
      if (nrm > 0.5) {
        for (int l = tid; l < ell; l += blockDim.x) Zout[int64_t(l) * k + t] = v[l] / nrm;
        __syncthreads();
        break;
      }
    }
  }
}

void svd_finalize(hipStream_t st, double* base, const SvdSector* secs, int nsec) {
  if (nsec <= 0) return;
  hipLaunchKernelGGL(k_svd_finalize, dim3(nsec), dim3(256), 0, st, base, secs);
}

__global__ __launch_bounds__(256) void k_svd_shift(double* __restrict__ base, const SvdSector* __restrict__ secs) {
  __shared__ double sh[256];
  const SvdSector s = secs[blockIdx.x];
  double* Y = base + s.y_off;
  const int64_t n = int64_t(s.ell) * s.k;
  double acc = 0.0;
  for (int64_t e = threadIdx.x; e < n; e += blockDim.x) acc += Y[e] * Y[e];
  const double c = sqrt(wg_sum(acc, sh));
  for (int i = threadIdx.x; i < s.k; i += blockDim.x) Y[int64_t(i) * s.ell + i] += c;
  if (threadIdx.x == 0) base[s.shift_off] = c;
}

void svd_shift(hipStream_t st, double* base, const SvdSector* secs, int nsec) {
  if (nsec <= 0) return;
  hipLaunchKernelGGL(k_svd_shift, dim3(nsec), dim3(256), 0, st, base, secs);
}

__global__ void k_fill_i32(int32_t* p, int n, int32_t v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

void fill_i32(hipStream_t st, int32_t* p, int n, int32_t v) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_i32, dim3((n + 255) / 256), dim3(256), 0, st, p, n, v);
}

// ---- global truncation ------------------------------------------------------
__global__ __launch_bounds__(256) void k_truncate(double* __restrict__ base, const int64_t* __restrict__ s_offs,
                                                  const int32_t* __restrict__ ks, int nsec, int64_t scratch_off,
                                                  int64_t total, double tol, double pw, int64_t min_size,
                                                  int64_t max_size, int32_t* __restrict__ kept,
                                                  double* __restrict__ thr_out) {
  __shared__ double thr_sh;
  double* v = base + scratch_off;
  int64_t n2 = 1;
  while (n2 < total) n2 <<= 1;
  const int tid = threadIdx.x;
  // gather (sector lists concatenated in order), pad with -inf
  {
    int64_t pos = 0;
    for (int i = 0; i < nsec; ++i) {
      const double* src = base + s_offs[i];
      for (int j = tid; j < ks[i]; j += blockDim.x) v[pos + j] = src[j];
      pos += ks[i];
    }
    for (int64_t j = total + tid; j < n2; j += blockDim.x) v[j] = -INFINITY;
  }
  __syncthreads();
  // bitonic sort, descending
  for (int64_t size = 2; size <= n2; size <<= 1) {
    for (int64_t stride = size >> 1; stride > 0; stride >>= 1) {
      for (int64_t i = tid; i < n2 / 2; i += blockDim.x) {
        const int64_t lo = 2 * stride * (i / stride) + (i % stride);
        const int64_t hi = lo + stride;
        const bool desc = ((lo & size) == 0);
        const double a = v[lo], b = v[hi];
        if (desc ? (a < b) : (a > b)) {
          v[lo] = b;
          v[hi] = a;
        }
      }
      __syncthreads();
    }
  }
  // compute_last_index (sources/LinearAlgebra.cpp:57-75), sequential as in the reference
  if (tid == 0) {
    const bool sq = (pw == 2.0);
    const double toln = sq ? tol * tol : pow(tol, pw);
    int64_t last = total - 1;
    double x = fabs(v[last]);
    double trunc = sq ? x * x : pow(x, pw);
    while (last >= min_size && last > 0) {
      if (trunc > toln && last < max_size) break;
      --last;
      x = fabs(v[last]);
      trunc += sq ? x * x : pow(x, pw);
    }
    const double sl = v[last];
    thr_sh = sl - 2.0 * sl * DBL_EPSILON;
    thr_out[0] = thr_sh;
  }
  __syncthreads();
  const double thr = thr_sh;
  for (int i = tid; i < nsec; i += blockDim.x) {
    const double* src = base + s_offs[i];
    int c = 0;
    while (c < ks[i] && src[c] > thr) ++c;
    kept[i] = c;
  }
}

void truncate_select(hipStream_t st, double* base, const int64_t* s_offs, const int32_t* ks, int nsec,
                     int64_t scratch_off, int64_t total, double tol, double pw, int64_t min_size, int64_t max_size,
                     int32_t* kept, double* thr_out) {
  hipLaunchKernelGGL(k_truncate, dim3(1), dim3(256), 0, st, base, s_offs, ks, nsec, scratch_off, total, tol, pw,
                     min_size, max_size, kept, thr_out);
}

}  // namespace kern
}  // namespace qtt
