// Gram-matrix eigenvectors for sectors wider than one work-group's LDS holds
// (k > kGramEigMaxK).  G = Y^T Y (k x k, column-major, leading dimension k)
// lives in device memory; the steps mirror the LDS solver of svd_jacobi.hip:
//
//   k_gram_check    orthogonality test of the columns behind G (relative, or
//                   absolute with respect to the largest singular value),
//   k_pchol_panel   one panel of nb columns of the pivoted Cholesky
//                   P^T G P = R^T R (left-looking inside the panel; the
//                   trailing matrix is updated by a GEMM G -= Lp Lp^T after
//                   each panel).  L = R^T is kept with its rows in the
//                   original order, so L L^T = G and the one-sided Jacobi on
//                   the columns of L (block rounds of k_svd_round) yields G's
//                   eigenvectors as the normalised columns,
//   k_gram_vectors  normalises the columns of the rotated L into W and
//                   completes zero columns from the unit vector with the
//                   largest residual (two projection passes).
#include <algorithm>
#include <cfloat>
#include <climits>
#include <stdexcept>

#include "../kernels.hpp"

namespace qtt {
namespace kern {

namespace {

__device__ double block_max(double v, double* sh) {
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (tid < w) sh[tid] = fmax(sh[tid], sh[tid + w]);
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

__device__ double block_sum(double v, double* sh) {
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

// 1/sqrt(x) by the hardware estimate and two Newton steps.
__device__ __forceinline__ double fast_rsqrt_pc(double x) {
  double y = __builtin_amdgcn_rsq(x);
  y = y * (1.5 - 0.5 * x * y * y);
  return y * (1.5 - 0.5 * x * y * y);
}

}  // namespace

constexpr int kLargeThreads = 1024;

__global__ void k_gram_check_zero(const SmallJacobi* __restrict__ js, int n, int* __restrict__ flags) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const SmallJacobi J = js[i];
    if (J.flag_idx >= 0) flags[J.flag_idx] = 0;
    if (J.aux_idx >= 0) flags[J.aux_idx] = 0;
  }
}

// flags[flag_idx] += pairs failing both tests, flags[aux_idx] += pairs passing
// only the absolute one (same tests as k_gram_eig, on squares), both zeroed
// beforehand.  Work-group y of a sector takes the columns c = y (mod
// gridDim.y); the diagonal sits in LDS (dynamic, k doubles); wave w scans the
// upper part of its w-th, (w + 16)-th, ... column with coalesced reads.
__global__ __launch_bounds__(kLargeThreads) void k_gram_check(const double* __restrict__ base,
                                                              const SmallJacobi* __restrict__ js,
                                                              int* __restrict__ flags, double tol) {
  extern __shared__ double dg[];
  __shared__ double sh[kLargeThreads];
  __shared__ int nv, na;
  const SmallJacobi J = js[blockIdx.x];
  const int k = J.k;
  const double* G = base + J.y_off;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwave = blockDim.x >> 6;
  const int ny = gridDim.y, y = blockIdx.y;
  if (tid == 0) nv = na = 0;
  double dm = 0.0;
  for (int i = tid; i < k; i += blockDim.x) {
    dg[i] = G[int64_t(i) * J.ldy + i];
    dm = fmax(dm, dg[i]);
  }
  const double g2max = block_max(dm, sh);  // also orders the dg writes before the reads
  const double t2 = tol * tol;
  int my = 0, mya = 0;
  for (int c = y + ny * wave; c < k; c += ny * nwave) {
    const double gjj = dg[c];
    const double* col = G + int64_t(c) * J.ldy;
#pragma unroll 4
    for (int r = lane; r < c; r += 64) {
      const double gii = dg[r], g = col[r];
      const double g2 = g * g;
      if (gii > 0.0 && gjj > 0.0 && g2 > t2 * gii * gjj) {
        if (g2 > t2 * fmax(gii, gjj) * g2max)
          ++my;
        else
          ++mya;
      }
    }
  }
  if (my) atomicAdd(&nv, my);
  if (mya) atomicAdd(&na, mya);
  __syncthreads();
  if (tid == 0) {
    if (J.flag_idx >= 0 && nv) atomicAdd(&flags[J.flag_idx], nv);
    if (J.aux_idx >= 0 && na) atomicAdd(&flags[J.aux_idx], na);
  }
}

void gram_check(hipStream_t st, const double* base, const void* descs, int n, int* flags, double tol, int kmax) {
  if (n <= 0) return;
  static const size_t limit = [] {
    int dev = 0, v = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) v = 65536;
    const size_t l = size_t(v) - kLargeThreads * sizeof(double) - 64;
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_gram_check),
                              hipFuncAttributeMaxDynamicSharedMemorySize, int(l));
    return l;
  }();
  const size_t lds = size_t(kmax) * sizeof(double);
  if (lds > limit) throw std::runtime_error("gram check: sector wider than the LDS holds");
  const int ny = std::max(1, std::min(8, kmax / 64));  // about 64 columns per work-group
  hipLaunchKernelGGL(k_gram_check_zero, dim3(1), dim3(256), 0, st, static_cast<const SmallJacobi*>(descs), n, flags);
  hipLaunchKernelGGL(k_gram_check, dim3(n, ny), dim3(kLargeThreads), lds, st, base,
                     static_cast<const SmallJacobi*>(descs), flags, tol);
}

// One panel [j0, j0 + nb) of the pivoted Cholesky of G (already updated by
// the previous panels).  state[0] = rank found so far (k while running),
// state[1 + i] = 1 once original index i has been pivoted.  L is k x ldl
// (column-major, leading dimension k), rows in the original order.  The
// panel's columns are also written to Lp (k x nb, column-major) for the
// trailing update.  LDS: the nb panel columns, the running diagonal dg
// (G_ii less this panel's columns so far) and the pivoted flags, k each.
// A step: per-thread best key over its indices, wave maxima by DPP and lane
// swizzles, one barrier, every thread reduces the wave maxima itself.  The
// key is the value's top 51 bits with the inverted index in the low 13
// (k <= 8192): the largest diagonal wins, the smallest index among values
// equal to 2^-39 relative; non-positive values never beat positive ones.
__device__ __forceinline__ long long pc_dpp(long long key, int ctrl) {
  int lo = int(key & 0xffffffffLL), hi = int(key >> 32);
  switch (ctrl) {
    case 0:
      lo = __builtin_amdgcn_mov_dpp(lo, 0xB1, 0xF, 0xF, false);
      hi = __builtin_amdgcn_mov_dpp(hi, 0xB1, 0xF, 0xF, false);
      break;
    case 1:
      lo = __builtin_amdgcn_mov_dpp(lo, 0x4E, 0xF, 0xF, false);
      hi = __builtin_amdgcn_mov_dpp(hi, 0x4E, 0xF, 0xF, false);
      break;
    case 2:
      lo = __builtin_amdgcn_mov_dpp(lo, 0x141, 0xF, 0xF, false);
      hi = __builtin_amdgcn_mov_dpp(hi, 0x141, 0xF, 0xF, false);
      break;
    case 3:
      lo = __builtin_amdgcn_mov_dpp(lo, 0x140, 0xF, 0xF, false);
      hi = __builtin_amdgcn_mov_dpp(hi, 0x140, 0xF, 0xF, false);
      break;
    case 4:  // lane xor 16 (ds_swizzle bit-mask mode)
      lo = __builtin_amdgcn_ds_swizzle(lo, 0x401F);
      hi = __builtin_amdgcn_ds_swizzle(hi, 0x401F);
      break;
    default:  // lane xor 32
      lo = __shfl_xor(lo, 32);
      hi = __shfl_xor(hi, 32);
      break;
  }
  return (long long)(unsigned)lo | ((long long)hi << 32);
}

__global__ __launch_bounds__(kLargeThreads) void k_pchol_panel(double* __restrict__ base,
                                                               const CholPanel* __restrict__ ps, int j0, int nb) {
  extern __shared__ double lds[];
  __shared__ long long wkey[2][kLargeThreads / 64];  // wave maxima, by step parity
  const CholPanel P = ps[blockIdx.x];
  const int k = P.k;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const double* G = base + P.g_off;
  double* L = base + P.l_off;
  double* Lp = base + P.lp_off;
  int* state = reinterpret_cast<int*>(base + P.state_off);
  double* col = lds;       // nb columns of length k
  double* dg = col + int64_t(nb) * k;
  int* used = reinterpret_cast<int*>(dg + k);
  if (state[0] < k || j0 >= k) {  // already rank-deficient or finished: zero panel
    for (int64_t e = tid; e < int64_t(k) * nb; e += blockDim.x) Lp[e] = 0.0;
    return;
  }
  // the key of index i: the diagonal's top bits with the inverted index below
  auto key_of = [&](int i) {
    return used[i] ? LLONG_MIN : ((__double_as_longlong(dg[i]) & ~0x1FFFLL) | (8191 - i));
  };
  auto publish = [&](long long key, int slot) {
#pragma unroll
    for (int st = 0; st < 6; ++st) {
      const long long o = pc_dpp(key, st);
      key = o > key ? o : key;
    }
    if (lane == 0) wkey[slot][wave] = key;
  };
  {
    long long key = LLONG_MIN;
    for (int i = tid; i < k; i += blockDim.x) {
      dg[i] = G[int64_t(i) * k + i];
      used[i] = state[1 + i];
      const long long kk = key_of(i);
      key = kk > key ? kk : key;
    }
    publish(key, 0);
  }
  __syncthreads();
  // One barrier per step: a thread forms the next step's key from its own
  // diagonal entries right after updating them, and the wave maxima go to
  // the other parity's slots (last read before the previous barrier).
  int rank = k;
  for (int jj = 0; jj < nb; ++jj) {
    const int j = j0 + jj;
    if (j >= k) {
      for (int i = tid; i < k; i += blockDim.x) col[jj * k + i] = 0.0;
      continue;
    }
    long long key = LLONG_MIN;
    // the launch has kLargeThreads threads: a fixed trip count lets the
    // sixteen reads be issued together
#pragma unroll
    for (int w = 0; w < kLargeThreads / 64; ++w) key = wkey[jj & 1][w] > key ? wkey[jj & 1][w] : key;
    const int p = key == LLONG_MIN ? -1 : 8191 - int(key & 0x1FFF);
    const double pv = p >= 0 ? dg[p] : 0.0;
    if (!(pv > 0.0)) {  // uniform: the rest is zero or negative noise
      for (int t = jj; t < nb; ++t)
        for (int i = tid; i < k; i += blockDim.x) col[t * k + i] = 0.0;
      rank = j;
      break;
    }
    const double isq = fast_rsqrt_pc(pv), ljj = pv * isq;
    long long nkey = LLONG_MIN;
    for (int i = tid; i < k; i += blockDim.x) {
      double v = 0.0;
      if (i == p) {
        v = ljj;
        used[p] = 1;
      } else if (!used[i]) {
        // the load of G's column p is independent of the dot product over the
        // panel's earlier columns, so its latency overlaps that sum
        const double gpi = G[int64_t(p) * k + i];
        double acc = 0.0;
        // unrolled so that the LDS reads of several terms are in flight at
        // once; the FMA chain keeps its order
#pragma unroll 8
        for (int t = 0; t < jj; ++t) acc = fma(col[t * k + i], col[t * k + p], acc);
        v = (gpi - acc) * isq;
        dg[i] -= v * v;
      }
      col[jj * k + i] = v;
      const long long kk = key_of(i);
      nkey = kk > nkey ? kk : nkey;
    }
    publish(nkey, (jj + 1) & 1);
    __syncthreads();
  }
  __syncthreads();  // the copies below read other threads' column entries
  for (int i = tid; i < k; i += blockDim.x) state[1 + i] = used[i];
  if (tid == 0 && rank < k) state[0] = rank;
  for (int64_t e = tid; e < int64_t(k) * nb; e += blockDim.x) {
    const int t = int(e / k), i = int(e % k);
    Lp[e] = col[e];
    if (j0 + t < k) L[int64_t(j0 + t) * k + i] = col[e];
  }
}

__global__ __launch_bounds__(256) void k_pchol_init(double* __restrict__ base, const CholPanel* __restrict__ ps) {
  const CholPanel P = ps[blockIdx.x];
  int* state = reinterpret_cast<int*>(base + P.state_off);
  for (int i = threadIdx.x; i <= P.k; i += blockDim.x) state[i] = i == 0 ? P.k : 0;
}

void pchol_init(hipStream_t st, double* base, const void* descs, int n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_pchol_init, dim3(n), dim3(256), 0, st, base, static_cast<const CholPanel*>(descs));
}

void pchol_panel(hipStream_t st, double* base, const void* descs, int n, int j0, int nb, int kmax) {
  if (n <= 0) return;
  const size_t lds = size_t(kmax) * (nb + kPcholExtra) * sizeof(double);
  hipLaunchKernelGGL(k_pchol_panel, dim3(n), dim3(kLargeThreads), lds, st, base,
                     static_cast<const CholPanel*>(descs), j0, nb);
}

size_t pchol_lds_limit() {
  static const size_t limit = [] {
    int dev = 0, v = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) v = 65536;
    const size_t l = size_t(v) - 1024;  // static arrays of the kernel
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_pchol_panel),
                              hipFuncAttributeMaxDynamicSharedMemorySize, int(l));
    return l;
  }();
  return limit;
}

// W = columns of X (k x k at x_off, leading dimension ldx) normalised, by
// k_gram_vec_norm over (sector, 64-column) work-groups (wave w takes four
// columns, rows in flight four at a time); then k_gram_vectors completes the
// zero columns from the unit vector with the largest residual (two
// projection passes), one work-group per sector, and returns at once when
// there are none.  W is written at w_off with leading dimension ldw;
// scratch_off: 3k doubles (norms, two work vectors).
__global__ __launch_bounds__(kLargeThreads) void k_gram_vec_norm(double* __restrict__ base,
                                                                 const GramVec* __restrict__ js) {
  const GramVec J = js[blockIdx.x];
  const int k = J.k;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c0 = blockIdx.y * 64 + wave * 4;
  if (c0 >= k) return;
  const double* X = base + J.x_off;
  double* W = base + J.w_off;
  double* nrm = base + J.scratch_off;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  const double* col[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) col[c] = X + int64_t(min(c0 + c, k - 1)) * J.ldx;
  for (int r0 = lane; r0 < k; r0 += 256) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int r = r0 + 64 * u;
      const int rr = r < k ? r : k - 1;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const double x = col[c][rr];
        acc[c] += r < k ? x * x : 0.0;
      }
    }
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    double a = acc[c];
    for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off);
    acc[c] = sqrt(a);
    if (lane == 0 && c0 + c < k) nrm[c0 + c] = acc[c];
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    if (c0 + c >= k) break;
    const double v = acc[c];
    double* w = W + int64_t(c0 + c) * J.ldw;
    for (int r = lane; r < k; r += 64) w[r] = v > 1e-300 ? col[c][r] / v : 0.0;
  }
}

__global__ __launch_bounds__(kLargeThreads) void k_gram_vectors(double* __restrict__ base,
                                                                const GramVec* __restrict__ js) {
  __shared__ double sh[kLargeThreads];
  __shared__ int best, nzero;
  const GramVec J = js[blockIdx.x];
  const int k = J.k;
  const int tid = threadIdx.x;
  double* W = base + J.w_off;
  double* nrm = base + J.scratch_off;  // k
  double* vec = nrm + k;               // k
  double* hh = vec + k;                // k
  if (tid == 0) nzero = 0;
  __syncthreads();
  int z = 0;
  for (int c = tid; c < k; c += blockDim.x) z += nrm[c] > 1e-300 ? 0 : 1;
  if (z) atomicAdd(&nzero, z);
  __syncthreads();
  if (nzero == 0) return;  // uniform
  for (int c = 0; c < k; ++c) {
    if (nrm[c] > 1e-300) continue;  // uniform
    for (int r = tid; r < k; r += blockDim.x) {
      double q = 1.0;
      for (int j = 0; j < k; ++j) q -= W[int64_t(j) * J.ldw + r] * W[int64_t(j) * J.ldw + r];
      vec[r] = q;
    }
    __syncthreads();
    if (tid == 0) {
      int bi = 0;
      for (int r = 1; r < k; ++r)
        if (vec[r] > vec[bi]) bi = r;
      best = bi;
    }
    __syncthreads();
    const int e0 = best;
    for (int r = tid; r < k; r += blockDim.x) {
      double q = r == e0 ? 1.0 : 0.0;
      for (int j = 0; j < k; ++j) q -= W[int64_t(j) * J.ldw + r] * W[int64_t(j) * J.ldw + e0];
      vec[r] = q;
    }
    __syncthreads();
    for (int j = tid; j < k; j += blockDim.x) {
      double q = 0.0;
      for (int r = 0; r < k; ++r) q += W[int64_t(j) * J.ldw + r] * vec[r];
      hh[j] = q;
    }
    __syncthreads();
    for (int r = tid; r < k; r += blockDim.x) {
      double q = vec[r];
      for (int j = 0; j < k; ++j) q -= W[int64_t(j) * J.ldw + r] * hh[j];
      vec[r] = q;
    }
    __syncthreads();
    double a = 0.0;
    for (int r = tid; r < k; r += blockDim.x) a += vec[r] * vec[r];
    const double n = sqrt(block_sum(a, sh));
    for (int r = tid; r < k; r += blockDim.x) W[int64_t(c) * J.ldw + r] = vec[r] / n;
    if (tid == 0) nrm[c] = n;
    __syncthreads();
  }
}

void gram_vectors(hipStream_t st, double* base, const void* descs, int n, int kmax) {
  if (n <= 0 || kmax <= 0) return;
  hipLaunchKernelGGL(k_gram_vec_norm, dim3(n, (kmax + 63) / 64), dim3(kLargeThreads), 0, st, base,
                     static_cast<const GramVec*>(descs));
  hipLaunchKernelGGL(k_gram_vectors, dim3(n), dim3(kLargeThreads), 0, st, base,
                     static_cast<const GramVec*>(descs));
}

}  // namespace kern
}  // namespace qtt
