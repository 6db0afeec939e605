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
#include <algorithm>
#include <cfloat>
#include <climits>
#include <stdexcept>

#include "../kernels.hpp"

namespace qtt {
namespace kern {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int NB2 = 2 * kSvdBlock;  // 32 columns per work-group
constexpr int YC = 128;             // rows of Y staged per chunk
constexpr int YP = YC + 2;          // LDS pitch of the staged Y chunk
constexpr int GP = NB2 + 1;         // LDS pitch of G and Q

// W = I (kp x kp) for every sector; gridDim.y work-groups share a sector,
// one column per wave-step.
__global__ void k_svd_init_w(double* __restrict__ base, const SvdSector* __restrict__ secs) {
  const SvdSector s = secs[blockIdx.x];
  const int kp = s.nb * kSvdBlock;
  double* W = base + s.w_off;
  const int wave = blockIdx.y * (blockDim.x >> 6) + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int nw = gridDim.y * (blockDim.x >> 6);
  for (int c = wave; c < kp; c += nw)
    for (int r = lane; r < kp; r += 64) W[int64_t(c) * kp + r] = (r == c) ? 1.0 : 0.0;
}

void svd_init_w(hipStream_t st, double* base, const SvdSector* secs, int nsec) {
  if (nsec <= 0) return;
  hipLaunchKernelGGL(k_svd_init_w, dim3(nsec, 16), dim3(256), 0, st, base, secs);
}

__device__ __forceinline__ int circle_block(int pos, int round, int nb) {
  return pos == 0 ? 0 : 1 + ((pos - 1 + round) % (nb - 1));
}

// Reciprocal square root and reciprocal by the hardware estimates refined by
// two Newton steps (each doubles the correct bits), well inside the rounding
// of the rotations they feed; the division and square root the compiler
// expands to correctly rounded sequences cost several times more latency.
__device__ __forceinline__ double fast_rsqrt(double x) {
  double y = __builtin_amdgcn_rsq(x);
  y = y * (1.5 - 0.5 * x * y * y);
  return y * (1.5 - 0.5 * x * y * y);
}
__device__ __forceinline__ double fast_rcp(double x) {
  double z = __builtin_amdgcn_rcp(x);
  z = z * (2.0 - x * z);
  return z * (2.0 - x * z);
}

// Jacobi rotation (c, s) annihilating the off-diagonal h/2 of the symmetric
// 2x2 [[a, h/2], [h/2, b]] with |angle| <= pi/4, from d = b - a:
// t = sign(d) h / (|d| + sqrt(d^2 + h^2)), c = 1/sqrt(1 + t^2), s = c t.
__device__ __forceinline__ void jacobi_cs(double d, double h, double& c, double& sn) {
  const double q2 = d * d + h * h;
  const double r = q2 * fast_rsqrt(q2);
  const double t = (d >= 0.0 ? h : -h) * fast_rcp(fabs(d) + r);
  c = fast_rsqrt(1.0 + t * t);
  sn = c * t;
}

// lane i of each 16-lane row <- lane (i + 1) mod 16 (DPP row rotate right by
// 15), on a double as two 32-bit moves.
__device__ __forceinline__ double dpp_row_next(double v) {
  const long long b = __double_as_longlong(v);
  int lo = int(b & 0xffffffffLL), hi = int(b >> 32);
  lo = __builtin_amdgcn_mov_dpp(lo, 0x12F, 0xF, 0xF, false);
  hi = __builtin_amdgcn_mov_dpp(hi, 0x12F, 0xF, 0xF, false);
  return __longlong_as_double((long long)(unsigned)lo | ((long long)hi << 32));
}

// Apply the 32x32 rotation Q (LDS, pitch GP) to the 32 selected columns of a
// column-major matrix M (rows x ld) in place, YC rows at a time (each wave
// owns two 16-row MFMA tiles of a chunk).
__device__ void apply_rotation(double* __restrict__ M, int64_t ld, int rows, const int* cols,
                               double (*Ys)[YP], const double (*Q)[GP]) {
  constexpr int PER = YC / 8;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const bool act = tid < 256;  // the first four waves work, every thread meets the barriers
  const int c = act ? tid >> 3 : 0, lq = (tid & 7) * PER;
  const double* src = M + int64_t(cols[c]) * ld;
  double buf[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i) buf[i] = (act && lq + i < rows) ? src[lq + i] : 0.0;
  for (int l0 = 0; l0 < rows; l0 += YC) {
    if (act) {
#pragma unroll
      for (int i = 0; i < PER; ++i) Ys[c][lq + i] = buf[i];
    }
    __syncthreads();
    if (act && l0 + YC < rows) {
#pragma unroll
      for (int i = 0; i < PER; ++i) {
        const int l = l0 + YC + lq + i;
        buf[i] = (l < rows) ? src[l] : 0.0;
      }
    }
    v4d acc[2][2] = {{v4d{0.0, 0.0, 0.0, 0.0}, v4d{0.0, 0.0, 0.0, 0.0}},
                     {v4d{0.0, 0.0, 0.0, 0.0}, v4d{0.0, 0.0, 0.0, 0.0}}};
#pragma unroll
    for (int kk = 0; kk < NB2 && act; kk += 4) {
      const int kr = kk + (lane >> 4), cl = lane & 15;
      const double a0 = Ys[kr][wave * 32 + cl], a1 = Ys[kr][wave * 32 + 16 + cl];
      const double b0 = Q[kr][cl], b1 = Q[kr][16 + cl];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
    __syncthreads();  // every wave has read its Ys rows before they are overwritten
    if (!act) continue;
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
      for (int jt = 0; jt < 2; ++jt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = l0 + wave * 32 + rt * 16 + 4 * r + (lane >> 4);  // f64 MFMA output layout
          const int col = jt * 16 + (lane & 15);
          if (row < rows) M[int64_t(cols[col]) * ld + row] = acc[rt][jt][r];
        }
  }
}

// RESIDENT: the block pair's ell x 32 columns are loaded into dynamic LDS
// once (ell <= the LDS capacity), the Gram matrix and the rotated columns are
// computed from there and written back once; otherwise 128-row chunks are
// staged twice (Gram, then rotation).
template <bool RESIDENT>
__global__ __launch_bounds__(RESIDENT ? 1024 : 256) void k_svd_round(double* __restrict__ base,
                                                   const RoundWork* __restrict__ work, int round,
                                                   int* __restrict__ counters, const int* __restrict__ skip,
                                                   double tol, int inner_sweeps, int pitch, int cross_after0) {
  extern __shared__ double dyn[];
  __shared__ double G[1][NB2][GP];
  __shared__ double Q[1][NB2][GP];
  __shared__ int cols[NB2];
  __shared__ int flag, nrot;
  double (*Ys)[YP] = reinterpret_cast<double (*)[YP]>(dyn);  // chunk staging (aliases the resident block)

  const RoundWork s = work[blockIdx.x];
  const int pi = s.pair;
  const int nb = s.nb;
  if (round >= nb - 1) return;
  // a sector whose previous sweep rotated nothing has converged
  if (skip && skip[s.flag_idx] == 0) return;
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
  if (RESIDENT) {
    // all 32 columns at once: 32 lanes per column take rows (tid & 31) + 32u,
    // up to 36 loads per thread in flight (one pass for the resident sizes)
    double* Yl = dyn;  // Yl[c * pitch + r]
    {
      constexpr int CH = 36;
      const int c = tid >> 5, part = tid & 31;
      const double* src = Y + int64_t(cols[c]) * ld;
      double* dst = Yl + c * pitch;
      for (int o = 0; o < ell; o += 32 * CH) {
        double v[CH];
        // 32-row slices wholly past ell are skipped (a uniform branch): their
        // clamped loads would only occupy the address units
#pragma unroll
        for (int u = 0; u < CH; ++u) {
          const int r = o + part + 32 * u;
          v[u] = 0.0;
          if (o + 32 * u < ell) v[u] = src[r < ell ? r : ell - 1];
        }
#pragma unroll
        for (int u = 0; u < CH; ++u) {
          const int r = o + part + 32 * u;
          if (r < ell) dst[r] = v[u];
        }
      }
    }
    __syncthreads();
    // G is symmetric: the three distinct 16x16 tiles (0,0), (0,1), (1,1),
    // five waves each on the 4-row slices 4 part + 20 t (two accumulation
    // chains); the fifths are summed into G in a fixed order below and the
    // (0,1) tile is mirrored into (1,0).  Wave 15 only meets the barriers.
    const int cl = lane & 15, kq = lane >> 4;
    constexpr int kParts = 5;
    const bool gw = wave < 3 * kParts;
    const int tile = gw ? wave % 3 : 0, part = gw ? wave / 3 : kParts;
    const int ti = tile == 2 ? 1 : 0, tj = tile == 0 ? 0 : 1;
    const double* ya = Yl + (ti * 16 + cl) * pitch + kq;
    const double* yb = Yl + (tj * 16 + cl) * pitch + kq;
    v4d g1 = v4d{0.0, 0.0, 0.0, 0.0};
    if (gw && ti == tj) {  // diagonal tile: both operands are the same columns
      int kk = 4 * part;
      for (; kk + 4 * kParts < ell; kk += 8 * kParts) {
        const bool in1 = kk + 4 * kParts + kq < ell;
        const double a0 = ya[kk];
        const double a1 = in1 ? ya[kk + 4 * kParts] : 0.0;
        g = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, a0, g, 0, 0, 0);
        g1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, a1, g1, 0, 0, 0);
      }
      for (; kk < ell; kk += 4 * kParts) {
        const double xa = kk + kq < ell ? ya[kk] : 0.0;
        g = __builtin_amdgcn_mfma_f64_16x16x4f64(xa, xa, g, 0, 0, 0);
      }
      g = g + g1;
    } else if (gw) {
      int kk = 4 * part;
      for (; kk + 4 * kParts < ell; kk += 8 * kParts) {
        const bool in1 = kk + 4 * kParts + kq < ell;
        const double a0 = ya[kk], b0 = yb[kk];
        const double a1 = ya[in1 ? kk + 4 * kParts : 0], b1 = yb[in1 ? kk + 4 * kParts : 0];
        g = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, g, 0, 0, 0);
        g1 = __builtin_amdgcn_mfma_f64_16x16x4f64(in1 ? a1 : 0.0, in1 ? b1 : 0.0, g1, 0, 0, 0);
      }
      for (; kk < ell; kk += 4 * kParts) {
        const bool in = kk + kq < ell;
        const double xa = ya[in ? kk : 0], xb = yb[in ? kk : 0];
        g = __builtin_amdgcn_mfma_f64_16x16x4f64(in ? xa : 0.0, in ? xb : 0.0, g, 0, 0, 0);
      }
      g = g + g1;
    }
#pragma unroll
    for (int qq = 0; qq < kParts; ++qq) {
      if (part == qq) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int gi = ti * 16 + 4 * r + (lane >> 4), gj = tj * 16 + (lane & 15);
          double& dst = G[0][gi][gj];
          const double v = qq == 0 ? g[r] : dst + g[r];
          dst = v;
          if (tile == 1) G[0][gj][gi] = v;
        }
      }
      __syncthreads();
    }
  } else {
    // YC-row chunks staged through LDS; the next chunk's global loads are
    // issued before the MFMAs of the current one
    constexpr int PER = YC / 8;
    const int c = tid >> 3, lq = (tid & 7) * PER;
    const double* src = Y + int64_t(cols[c]) * ld;
    double buf[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) buf[i] = (lq + i < ell) ? src[lq + i] : 0.0;
    for (int l0 = 0; l0 < ell; l0 += YC) {
#pragma unroll
      for (int i = 0; i < PER; ++i) Ys[c][lq + i] = buf[i];
      __syncthreads();
      if (l0 + YC < ell) {
#pragma unroll
        for (int i = 0; i < PER; ++i) {
          const int l = l0 + YC + lq + i;
          buf[i] = (l < ell) ? src[l] : 0.0;
        }
      }
#pragma unroll
      for (int kk = 0; kk < YC; kk += 4) {
        const int kr = kk + (lane >> 4), cl = lane & 15;
        g = __builtin_amdgcn_mfma_f64_16x16x4f64(Ys[wi * 16 + cl][kr], Ys[wj * 16 + cl][kr], g, 0, 0, 0);
      }
      __syncthreads();
    }
  }
  if (!RESIDENT) {  // the resident path has summed its quarters into G already
#pragma unroll
    for (int r = 0; r < 4; ++r) G[0][wi * 16 + 4 * r + (lane >> 4)][wj * 16 + (lane & 15)] = g[r];
  }
  __syncthreads();
  // ---- 2. convergence test ---------------------------------------------------
  for (int e = tid; e < NB2 * NB2; e += blockDim.x) {
    const int i = e / NB2, j = e % NB2;
    if (i < j) {
      const double gii = G[0][i][i], gjj = G[0][j][j], gij = G[0][i][j];
      if (gii > 0.0 && gjj > 0.0 && gij * gij > (tol * gii) * (tol * gjj)) flag = 1;
    }
  }
  __syncthreads();
  if (flag == 0) return;
  if (tid == 0) atomicAdd(&counters[s.flag_idx], 1);
  // ---- 3. eigen-decomposition of G by cyclic two-sided Jacobi -----------------
  // Thread (P, R) owns the 2x2 block of G at the row pair P and column pair R
  // of the round (circle ordering, 16 disjoint pairs): it computes both
  // pairs' rotations from the diagonal blocks, then applies J_P^T (rows) and
  // J_R (columns) to its block and J_R to the same rows of Q, in place.  The
  // 256 blocks tile G and Q, so one round is two barriers and no other LDS
  // traffic.
  for (int e = tid; e < NB2 * NB2; e += blockDim.x) {
    const int i = e / NB2, j = e % NB2;
    Q[0][i][j] = (i == j) ? 1.0 : 0.0;
  }
  __syncthreads();
  {
    // Step r: the 16 threads tid < 16 compute the rotations of the 16 pairs
    // (one each) into rc/rs; after a barrier thread (P, R) applies J_P^T
    // (rows) and J_R (columns) to its 2x2 block of G and J_R to the same rows
    // of Q, in place.  The 256 blocks tile G and Q, so a step is two barriers
    // and no chain of dependent LDS accesses beyond the diagonal reads.
    // With cross_after0, rounds after the first of a sweep take only the 256
    // pairs across the two blocks (pairs (i, 16 + (i + r) mod 16), 16
    // steps): the first round of every sweep holds each block once and takes
    // all 496 pairs, so a sweep still meets every column pair.
    __shared__ double rc[NB2 / 2], rs[NB2 / 2];
    const bool blk = tid < NB2 * NB2 / 4;  // 256 threads own the 2x2 blocks; the rest meet the barriers
    const int P = blk ? tid >> 4 : 0, R = tid & 15;
    const bool cross = cross_after0 && round > 0;
    const int nsteps = cross ? NB2 / 2 : NB2 - 1;
    auto pair_of = [&](int i, int r, int& x, int& y) {
      if (cross) {
        x = i;
        y = NB2 / 2 + ((i + r) & (NB2 / 2 - 1));
      } else {
        const int x0 = circle_block(i, r, NB2), x1 = circle_block(NB2 - 1 - i, r, NB2);
        x = min(x0, x1);
        y = max(x0, x1);
      }
    };
    if (cross && inner_sweeps == 1) {
      // The same cross sweep with G and Q held in registers.  Thread (P, R)
      // keeps g00 = G[P][R], g01 = G[P][b2], g10 = G[a2][R], g11 = G[a2][b2]
      // and Q's (P, R), (P, b2), (P+16, R), (P+16, b2), where a2 = 16 +
      // (P + r) mod 16 and b2 = 16 + (R + r) mod 16 move with the step r:
      // the diagonal owner (P == R) holds pair P's 2x2 block and computes
      // its rotation; after the update the moving columns come from lane
      // R + 1 of the same 16-lane row (DPP row rotate) and the moving rows
      // from thread P + 1 through LDS (G's area, which is dead from here
      // on).  After 16 steps every value is back at its first position.
      double (*s10)[NB2 / 2 + 1] = reinterpret_cast<double (*)[NB2 / 2 + 1]>(&G[0][0][0]);
      double (*s11)[NB2 / 2 + 1] = s10 + NB2 / 2;
      double g00 = 0.0, g01 = 0.0, g10 = 0.0, g11 = 0.0, qa0 = 0.0, qa1 = 0.0, qb0 = 0.0, qb1 = 0.0;
      constexpr int H = NB2 / 2;
      if (blk) {
        g00 = G[0][P][R];
        g01 = G[0][P][H + R];
        g10 = G[0][H + P][R];
        g11 = G[0][H + P][H + R];
        qa0 = Q[0][P][R];
        qa1 = Q[0][P][H + R];
        qb0 = Q[0][H + P][R];
        qb1 = Q[0][H + P][H + R];
      }
      __syncthreads();  // every read of G precedes the exchange writes into its area
      for (int r = 0; r < H; ++r) {
        if (blk && P == R) {
          const double app = g00, aqq = g11, apq = g01;
          const bool rot = app > 0.0 && aqq > 0.0 && apq * apq > (DBL_EPSILON * app) * (DBL_EPSILON * aqq);
          double cc, ss;
          jacobi_cs(aqq - app, 2.0 * apq, cc, ss);
          rc[P] = rot ? cc : 1.0;
          rs[P] = rot ? ss : 0.0;
        }
        __syncthreads();  // rotations published
        if (blk) {
          const double cp = rc[P], sp = rs[P], cq = rc[R], sq = rs[R];
          const double h00 = cp * g00 - sp * g10, h01 = cp * g01 - sp * g11;
          const double h10 = cp * g10 + sp * g00, h11 = cp * g11 + sp * g01;
          g00 = cq * h00 - sq * h01;
          g01 = cq * h01 + sq * h00;
          g10 = cq * h10 - sq * h11;
          g11 = cq * h11 + sq * h10;
          const double na1 = cq * qa1 + sq * qa0, nb1 = cq * qb1 + sq * qb0;
          qa0 = cq * qa0 - sq * qa1;
          qb0 = cq * qb0 - sq * qb1;
          g01 = dpp_row_next(g01);
          g11 = dpp_row_next(g11);
          qa1 = dpp_row_next(na1);
          qb1 = dpp_row_next(nb1);
          s10[P][R] = g10;
          s11[P][R] = g11;
        }
        __syncthreads();  // moving rows published; the next step's rotations wait for these reads
        if (blk) {
          g10 = s10[(P + 1) & (H - 1)][R];
          g11 = s11[(P + 1) & (H - 1)][R];
        }
      }
      __syncthreads();
      if (blk) {
        Q[0][P][R] = qa0;
        Q[0][P][H + R] = qa1;
        Q[0][H + P][R] = qb0;
        Q[0][H + P][H + R] = qb1;
      }
      __syncthreads();
    } else
    for (int sweep = 0; sweep < inner_sweeps; ++sweep) {
      if (tid == 0) nrot = 0;
      __syncthreads();
      for (int r = 0; r < nsteps; ++r) {
        int a, a2, b, b2;
        pair_of(P, r, a, a2);
        pair_of(R, r, b, b2);
        // The block owners issue their reads first, so that their latency
        // runs under the rotations below (only the owners read: the other
        // waves would multiply the LDS traffic).
        double g00 = 0.0, g01 = 0.0, g10 = 0.0, g11 = 0.0, q00 = 0.0, q01 = 0.0, q10 = 0.0, q11 = 0.0;
        if (blk) {
          g00 = G[0][a][b];
          g01 = G[0][a][b2];
          g10 = G[0][a2][b];
          g11 = G[0][a2][b2];
          q00 = Q[0][a][b];
          q01 = Q[0][a][b2];
          q10 = Q[0][a2][b];
          q11 = Q[0][a2][b2];
        }
        if (tid < NB2 / 2) {
          int x, y;
          pair_of(tid, r, x, y);
          const double app = G[0][x][x], aqq = G[0][y][y], apq = G[0][x][y];
          // computed unconditionally and selected, so that the chain has no
          // branch (an unselected result may be NaN, e.g. for a zero column)
          const bool rot = app > 0.0 && aqq > 0.0 && apq * apq > (DBL_EPSILON * app) * (DBL_EPSILON * aqq);
          double cc, ss;
          jacobi_cs(aqq - app, 2.0 * apq, cc, ss);
          rc[tid] = rot ? cc : 1.0;
          rs[tid] = rot ? ss : 0.0;
          if (rot) nrot = 1;  // a flag: every writer stores the same value
        }
        __syncthreads();  // rotations published; every read of this step precedes the writes
        if (blk) {
          const double cp = rc[P], sp = rs[P], cq = rc[R], sq = rs[R];
          // rows a, a2 by J_P^T; columns b, b2 by J_R (row' a = c a - s a2, row' a2 = c a2 + s a)
          const double h00 = cp * g00 - sp * g10, h01 = cp * g01 - sp * g11;
          const double h10 = cp * g10 + sp * g00, h11 = cp * g11 + sp * g01;
          G[0][a][b] = cq * h00 - sq * h01;
          G[0][a][b2] = cq * h01 + sq * h00;
          G[0][a2][b] = cq * h10 - sq * h11;
          G[0][a2][b2] = cq * h11 + sq * h10;
          Q[0][a][b] = cq * q00 - sq * q01;
          Q[0][a][b2] = cq * q01 + sq * q00;
          Q[0][a2][b] = cq * q10 - sq * q11;
          Q[0][a2][b2] = cq * q11 + sq * q10;
        }
        __syncthreads();
      }
      if (nrot == 0) break;
      __syncthreads();
    }
  }
  const int cur = 0;
  // ---- 4. rotate the columns of Y and W ---------------------------------------
  if (RESIDENT) {
    // (Y Q)^T = Q^T Y^T tile by tile: the MFMA output then holds 16
    // consecutive rows of one column per 16 lanes, so the stores coalesce
    const double* Yl = dyn;
    const int nw = blockDim.x >> 6;
    for (int l0 = 0; l0 < ell; l0 += nw * 32) {
      v4d acc[2][2] = {{v4d{0.0, 0.0, 0.0, 0.0}, v4d{0.0, 0.0, 0.0, 0.0}},
                       {v4d{0.0, 0.0, 0.0, 0.0}, v4d{0.0, 0.0, 0.0, 0.0}}};
      const int cl = lane & 15;
      const int r0 = l0 + wave * 32 + cl, r1 = r0 + 16;
#pragma unroll
      for (int kk = 0; kk < NB2; kk += 4) {
        const int kr = kk + (lane >> 4);
        const double a0 = r0 < ell ? Yl[kr * pitch + r0] : 0.0;
        const double a1 = r1 < ell ? Yl[kr * pitch + r1] : 0.0;
        const double b0 = Q[cur][kr][cl], b1 = Q[cur][kr][16 + cl];
        acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(b0, a0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(b1, a0, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(b0, a1, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(b1, a1, acc[1][1], 0, 0, 0);
      }
#pragma unroll
      for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int jt = 0; jt < 2; ++jt)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int row = (rt ? r1 : r0);
            const int col = jt * 16 + 4 * r + (lane >> 4);  // f64 MFMA output layout, transposed tile
            if (row < ell) Y[int64_t(cols[col]) * ld + row] = acc[rt][jt][r];
          }
    }
  } else {
    apply_rotation(Y, ld, ell, cols, Ys, Q[cur]);
  }
  if (s.w_off < 0) return;  // caller keeps no right rotation
  __syncthreads();           // the staging area below aliases the resident block
  const int kp = nb * kSvdBlock;
  apply_rotation(base + s.w_off, kp, kp, cols, Ys, Q[cur]);
}

// Rows of a block pair that stay resident in LDS (dynamic size at most the
// device's per-block LDS less the static arrays).
// (Thread-safe static initialisation: concurrent solves on several host
// threads see the limit only after the attribute has been set.)
int svd_round_dyn_max() {
  static const int dyn_max = [] {
    int dev = 0, v = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) v = 65536;
    const int m = v - int(2 * NB2 * GP * sizeof(double)) - 1024;
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_svd_round<true>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, m);
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_svd_round<false>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, m);
    return m;
  }();
  return dyn_max;
}

// LDS pitch (doubles) of a resident column: at least ell and 2 (mod 32), so
// that a 32-lane phase of the Gram's reads (16 columns x 2 rows) meets each
// of the 64 banks once.
int resident_pitch(int ell) { return (ell + 29) / 32 * 32 + 2; }

void svd_round(hipStream_t st, double* base, const RoundWork* work, int nwork, int round, int* counters,
               const int* skip, double tol, int inner_sweeps, int max_ell, int cross_after0) {
  if (nwork <= 0) return;
  const size_t staging = size_t(NB2) * YP * sizeof(double);
  const int pitch = resident_pitch(max_ell);
  if (int64_t(pitch) * NB2 * int64_t(sizeof(double)) <= svd_round_dyn_max()) {
    const size_t lds = std::max(size_t(NB2) * pitch * sizeof(double), staging);
    hipLaunchKernelGGL(k_svd_round<true>, dim3(nwork), dim3(1024), lds, st, base, work, round, counters, skip, tol,
                       inner_sweeps, pitch, cross_after0);
  } else {
    hipLaunchKernelGGL(k_svd_round<false>, dim3(nwork), dim3(256), staging, st, base, work, round, counters, skip,
                       tol, inner_sweeps, 0, cross_after0);
  }
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

// ---- single-work-group Jacobi for sectors that fit in LDS -------------------
// One work-group per sector holds Y (ell x ke, ke = k rounded up to even) in
// LDS and runs cyclic one-sided Jacobi sweeps (circle ordering over single
// columns) until a sweep rotates nothing.  Each column pair is handled by a
// group of gsz lanes that reduce the three dot products with shuffles and
// rotate the pair in LDS; W (ldw x ldw, column-major, global) accumulates the
// rotations.  Y is written back at the end; flags[flag_idx] receives the
// number of rotations of the first sweep.
constexpr int kSmallThreads = 1024;

__global__ __launch_bounds__(kSmallThreads) void k_jacobi_small(double* __restrict__ base,
                                                                const SmallJacobi* __restrict__ js,
                                                                int* __restrict__ flags, double tol, int max_sweeps) {
  extern __shared__ double lds[];
  __shared__ int nrot[2];
  __shared__ int first;
  const SmallJacobi J = js[blockIdx.x];
  const int ell = J.ell, k = J.k;
  const int ke = k + (k & 1);
  double* Yg = base + J.y_off;
  double* W = base + J.w_off;
  const int tid = threadIdx.x;
  // load Y (padding column zero) and set W to the identity
  for (int64_t e = tid; e < int64_t(ell) * ke; e += blockDim.x) {
    const int r = int(e % ell), c = int(e / ell);
    lds[e] = c < k ? Yg[int64_t(c) * J.ldy + r] : 0.0;
  }
  for (int64_t e = tid; e < int64_t(J.ldw) * J.ldw; e += blockDim.x)
    W[e] = (e % J.ldw == e / J.ldw) ? 1.0 : 0.0;
  if (tid == 0) {
    first = -1;
    nrot[0] = nrot[1] = 0;
  }
  __syncthreads();
  const int npairs = ke / 2;
  int gsz = 64;
  while (gsz > 1 && gsz * npairs > int(blockDim.x)) gsz >>= 1;
  const int pi = tid / gsz, li = tid % gsz, pstride = int(blockDim.x) / gsz;
  for (int sweep = 0; sweep < max_sweeps && ke >= 2; ++sweep) {
    int* cnt = &nrot[sweep & 1];
    for (int r = 0; r < ke - 1; ++r) {
      for (int p = pi; p < npairs; p += pstride) {
        const int a0 = circle_block(p, r, ke), b0 = circle_block(ke - 1 - p, r, ke);
        const int a = min(a0, b0), b = max(a0, b0);
        double* ya = lds + int64_t(a) * ell;
        double* yb = lds + int64_t(b) * ell;
        double al = 0.0, be = 0.0, ga = 0.0;
#pragma unroll 4
        for (int l = li; l < ell; l += gsz) {
          const double x = ya[l], y = yb[l];
          al += x * x;
          be += y * y;
          ga += x * y;
        }
        for (int off = gsz >> 1; off > 0; off >>= 1) {
          al += __shfl_xor(al, off, gsz);
          be += __shfl_xor(be, off, gsz);
          ga += __shfl_xor(ga, off, gsz);
        }
        if (al > 0.0 && be > 0.0 && fabs(ga) > tol * sqrt(al) * sqrt(be)) {
          const double zeta = (be - al) / (2.0 * ga);
          const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
          const double c = 1.0 / sqrt(1.0 + t * t), sn = c * t;
#pragma unroll 4
          for (int l = li; l < ell; l += gsz) {
            const double x = ya[l], y = yb[l];
            ya[l] = c * x - sn * y;
            yb[l] = sn * x + c * y;
          }
          double* wa = W + int64_t(a) * J.ldw;
          double* wb = W + int64_t(b) * J.ldw;
#pragma unroll 4
          for (int l = li; l < k; l += gsz) {
            const double x = wa[l], y = wb[l];
            wa[l] = c * x - sn * y;
            wb[l] = sn * x + c * y;
          }
          if (li == 0) atomicAdd(cnt, 1);
        }
      }
      __syncthreads();
    }
    // cnt is final after the last barrier; the other counter is cleared for
    // the next sweep before anyone can touch it (it is only used after the
    // next sweep's first barrier)
    const int n = *cnt;
    if (tid == 0) {
      if (first < 0) first = n;
      nrot[(sweep + 1) & 1] = 0;
    }
    __syncthreads();
    if (n == 0) break;
  }
  for (int64_t e = tid; e < int64_t(ell) * k; e += blockDim.x) {
    const int r = int(e % ell), c = int(e / ell);
    Yg[int64_t(c) * J.ldy + r] = lds[e];
  }
  if (tid == 0 && J.flag_idx >= 0) flags[J.flag_idx] = first < 0 ? 0 : first;
}

void jacobi_small(hipStream_t st, double* base, const void* descs, int n, int* flags, double tol, int max_sweeps,
                  size_t lds_bytes) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_jacobi_small, dim3(n), dim3(kSmallThreads), lds_bytes, st, base,
                     static_cast<const SmallJacobi*>(descs), flags, tol, max_sweeps);
}

// ---- eigenvectors of a Gram matrix by one-sided Jacobi without accumulation --
// One work-group per k x k Gram matrix G (k <= kGramEigMaxK) held in LDS.
//
// 1. Test: a pair (i, j) of the columns behind G is converged when
//    |g_ij| <= tol sqrt(g_ii g_jj) (relative orthogonality) or
//    |g_ij| <= tol sqrt(max(g_ii, g_jj)) sqrt(max_l g_ll) (its coupling moves
//    the larger column's share of the smaller one by at most tol times the
//    largest singular value: absolute accuracy, the class of LAPACK gesdd).
//    flags[flag_idx] = pairs failing both; flags[J.aux_idx] (when >= 0) = pairs
//    converged only in the absolute sense.  No failing pair: W = I exactly.
// 2. Preconditioning (precond != 0): pivoted Cholesky P^T G P = R^T R in
//    place (diagonal pivoting, stops at a non-positive pivot), X = R^T.  This
//    is the QR-with-pivoting preconditioner of Drmac and Veselic (SIAM J.
//    Matrix Anal. Appl. 29, 2008) applied through G: on graded spectra the
//    Jacobi below then needs a few sweeps instead of tens.  Without it X = G.
// 3. One-sided Jacobi drives X W to mutually orthogonal columns.  Since
//    X X^T = P^T G P (or G^2), the normalised columns of the result are the
//    eigenvectors of P^T G P and no rotation is accumulated.  Rotations are
//    made down to tol/16, so the columns end orthonormal well inside tol
//    whatever the accuracy of the small eigenpairs; zero columns are
//    completed from the unit vector with the largest residual (two
//    projection passes).
// 4. W is written with the rows un-permuted (W = P Q).
// Sum over aligned groups of g lanes (g a power of two <= 16) with DPP moves:
// xor 1 and xor 2 by quad permutes, then the 8- and 16-lane mirrors (after
// the previous steps every lane of a half holds the same partial sum, so the
// mirror partner holds the other half's).
__device__ __forceinline__ double dpp_bits(double v, int ctrl_id) {
  const long long b = __double_as_longlong(v);
  int lo = int(b & 0xffffffffLL), hi = int(b >> 32);
  switch (ctrl_id) {
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
    default:
      lo = __builtin_amdgcn_mov_dpp(lo, 0x140, 0xF, 0xF, false);
      hi = __builtin_amdgcn_mov_dpp(hi, 0x140, 0xF, 0xF, false);
      break;
  }
  return __longlong_as_double((long long)(unsigned)lo | ((long long)hi << 32));
}

__device__ __forceinline__ void group_sum3(double& a, double& b, double& c, int g) {
  if (g > 16) {
    for (int off = g >> 1; off > 0; off >>= 1) {
      a += __shfl_xor(a, off, g);
      b += __shfl_xor(b, off, g);
      c += __shfl_xor(c, off, g);
    }
    return;
  }
  for (int st = 0; (2 << st) <= g; ++st) {
    const double ta = dpp_bits(a, st), tb = dpp_bits(b, st), tc = dpp_bits(c, st);
    a += ta;
    b += tb;
    c += tc;
  }
}

// Lanes per column pair: enough for about 16 rows per lane, at most what the
// block holds for all pairs.
__host__ __device__ inline int gram_eig_gsz(int k, int threads) {
  const int npairs = (k + (k & 1)) / 2;
  int g = 1;
  while (g < 16 && g * 16 < k) g <<= 1;
  while (g > 1 && g * npairs > threads) g >>= 1;
  return g;
}


// ---- wave-local block Jacobi ---------------------------------------------------
// The columns of X form NB = 2*ceil(k/16) blocks of 8 (columns >= k are zero
// and never stored).  A work-group round pairs the blocks by the circle
// ordering, one block pair (P, Q) per wave.  Inside the wave, 8-lane group g
// keeps column g of P and column g of Q in registers (rows li + 8j), so the 64
// column pairs across P and Q take 8 steps with no barrier: between steps the
// Q columns move between groups by a lane xor (group offsets in Gray-code
// order 0,1,3,2,6,7,5,4, so every group meets every Q column once).  The first
// round of a sweep also takes the 28 pairs inside each block, by xor pairing
// of the groups (both groups of a pair compute the same rotation from the same
// sums and update their own column).  One barrier per round instead of one per
// column pair, and the columns cross LDS once per round.  Each lane keeps R
// rows of a column, R the smallest even count with 8 R >= k.
// Nine waves: one per block pair at k = 144 (18 blocks), and at most three
// waves per SIMD, so each keeps two 18-row columns and the shifted copy in
// registers without spilling.
constexpr int kGramEigThreads = 576;
static_assert(kGramEigMaxK <= 8 * 18, "the widest wave-local variant holds 18 rows per lane");

__device__ __forceinline__ double dpp_ror8(double v) {  // lane i <- lane i ^ 8 within 16-lane rows
  const long long b = __double_as_longlong(v);
  int lo = int(b & 0xffffffffLL), hi = int(b >> 32);
  lo = __builtin_amdgcn_mov_dpp(lo, 0x128, 0xF, 0xF, false);
  hi = __builtin_amdgcn_mov_dpp(hi, 0x128, 0xF, 0xF, false);
  return __longlong_as_double((long long)(unsigned)lo | ((long long)hi << 32));
}

// Lane permutations without index registers where the hardware allows:
// xor 8 by a DPP row rotation, xor 16 / 24 by ds_swizzle (bit-mask mode,
// and 0x1F, xor in bits 10-14), xor 32 by ds_bpermute.
template <int PATTERN>
__device__ __forceinline__ double swizzle64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_ds_swizzle(int(b & 0xffffffffLL), PATTERN);
  const int hi = __builtin_amdgcn_ds_swizzle(int(b >> 32), PATTERN);
  return __longlong_as_double((long long)(unsigned)lo | ((long long)hi << 32));
}
__device__ __forceinline__ double xor32(double v) { return __shfl_xor(v, 32); }

template <int M>
__device__ __forceinline__ double xor_groups(double v) {  // lane i <- lane i ^ (8 M)
  if constexpr (M == 1) return dpp_ror8(v);
  if constexpr (M == 2) return swizzle64<0x401F>(v);
  if constexpr (M == 3) return swizzle64<0x601F>(v);
  if constexpr (M == 4) return xor32(v);
  if constexpr (M == 5) return xor32(dpp_ror8(v));
  if constexpr (M == 6) return xor32(swizzle64<0x401F>(v));
  return xor32(swizzle64<0x601F>(v));
}

__device__ __forceinline__ void sum8(double& a) {
#pragma unroll
  for (int st = 0; st < 3; ++st) a += dpp_bits(a, st);
}

// A column in registers: true column = d * v (rows li + 8j), n = its squared
// norm.  Rotations in the scaled form move two FMAs per row (x' = x - t r y
// with r a ratio of scales) and fold the cosine into d; n follows the exact
// identities |x'|^2 = |x|^2 - t x.y, |y'|^2 = |y|^2 + t x.y and is recomputed
// from the rows when it falls below a quarter of its value (cancellation).
// The scales start at 1 when a column is loaded, so with at most 15
// rotations between a load and a store they stay above 0.7^15.
template <int R>
struct WlCol {
  double v[R];
  double d, n;
};

template <int R>
__device__ __forceinline__ double wl_norm(const WlCol<R>& x) {
  double s = 0.0;
#pragma unroll
  for (int j = 0; j < R; ++j) s += x.v[j] * x.v[j];
  sum8(s);
  return s * x.d * x.d;
}

// Column c of X (pitch k) into rows li + 8j.  Rows past k read the next
// column (or, for the last one, the arrays that follow X in LDS) and are
// masked; columns past k are zero.  One base address per column keeps the
// loads on immediate offsets.
template <int R>
__device__ __forceinline__ void wl_load(const double* X, int k, int c, int li, WlCol<R>& x) {
  const bool cin = c < k;
  const double* col = X + (cin ? c * k : 0) + li;
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const double v = col[8 * j];
    x.v[j] = (cin && li + 8 * j < k) ? v : 0.0;
  }
  x.d = 1.0;
  x.n = wl_norm<R>(x);
}

template <int R>
__device__ __forceinline__ void wl_store(double* X, int k, int c, int li, const WlCol<R>& x) {
  if (c >= k) return;
  double* col = X + c * k + li;
#pragma unroll
  for (int j = 0; j < R; ++j)
    if (li + 8 * j < k) col[8 * j] = x.v[j] * x.d;
}

// Rotation of the column pair (x, y) from their squared norms and product:
// t = tan(angle), c = cos(angle) as in jacobi_cs (x' = c x - s y,
// y' = s x + c y, s = c t); false when the pair is orthogonal to tol_rot.
__device__ __forceinline__ bool wl_rotation(double xx, double yy, double xy, double tol_rot, double& t, double& c) {
  if (xx > 0.0 && yy > 0.0 && xy * xy > (tol_rot * xx) * (tol_rot * yy)) {
    const double d = yy - xx, h = 2.0 * xy;
    const double q2 = d * d + h * h;
    const double r = q2 * fast_rsqrt(q2);
    t = (d >= 0.0 ? h : -h) * fast_rcp(fabs(d) + r);
    c = fast_rsqrt(1.0 + t * t);
    return true;
  }
  return false;
}

// Squared norm after a rotation, recomputed from the rows on cancellation.
template <int R>
__device__ __forceinline__ double wl_next_norm(const WlCol<R>& x, double n_old, double n_new) {
  if (n_new < 0.25 * n_old) n_new = wl_norm<R>(x);  // uniform in the group
  return n_new;
}

// Pairs (g, g ^ M) inside one block: x is the lower group's column.  Both
// groups form the same sums (the products commute and the reduction order is
// the same), so they compute the same rotation and each updates its own
// column.
template <int R, int M>
__device__ __forceinline__ void wl_intra(WlCol<R>& x, int grp, double tol_rot, int& cnt) {
  const bool lower = grp < (grp ^ M);
  double p[R];
#pragma unroll
  for (int j = 0; j < R; ++j) p[j] = xor_groups<M>(x.v[j]);
  const double pd = xor_groups<M>(x.d), pn = xor_groups<M>(x.n);
  double r = 0.0;
#pragma unroll
  for (int j = 0; j < R; ++j) r += x.v[j] * p[j];
  sum8(r);
  const double xy = (x.d * pd) * r;
  double t, c;
  if (wl_rotation(lower ? x.n : pn, lower ? pn : x.n, xy, tol_rot, t, c)) {
    // lower: x' = x - t (d_y / d_x) y;  upper: y' = y + t (d_x / d_y) x
    const double f = (lower ? -t : t) * pd * fast_rcp(x.d);
#pragma unroll
    for (int j = 0; j < R; ++j) x.v[j] += f * p[j];
    x.d *= c;
    x.n = wl_next_norm<R>(x, x.n, x.n + (lower ? -t : t) * xy);
    cnt += lower ? 1 : 0;
  }
}

template <int R>
__device__ __forceinline__ void wl_intra_all(WlCol<R>& x, int grp, double tol_rot, int& cnt) {
  wl_intra<R, 1>(x, grp, tol_rot, cnt);
  wl_intra<R, 2>(x, grp, tol_rot, cnt);
  wl_intra<R, 3>(x, grp, tol_rot, cnt);
  wl_intra<R, 4>(x, grp, tol_rot, cnt);
  wl_intra<R, 5>(x, grp, tol_rot, cnt);
  wl_intra<R, 6>(x, grp, tol_rot, cnt);
  wl_intra<R, 7>(x, grp, tol_rot, cnt);
}

template <int R>
__device__ __forceinline__ void wl_cross(WlCol<R>& a, WlCol<R>& b, double tol_rot, int& cnt) {
  double r = 0.0;
#pragma unroll
  for (int j = 0; j < R; ++j) r += a.v[j] * b.v[j];
  sum8(r);
  const double ab = (a.d * b.d) * r;
  double t, c;
  if (wl_rotation(a.n, b.n, ab, tol_rot, t, c)) {
    const double fa = -t * b.d * fast_rcp(a.d), fb = t * a.d * fast_rcp(b.d);
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const double x = a.v[j], y = b.v[j];
      a.v[j] = x + fa * y;
      b.v[j] = y + fb * x;
    }
    a.d *= c;
    b.d *= c;
    a.n = wl_next_norm<R>(a, a.n, a.n - t * ab);
    b.n = wl_next_norm<R>(b, b.n, b.n + t * ab);
    ++cnt;
  }
}

template <int R, int M>
__device__ __forceinline__ void wl_shift(WlCol<R>& b) {
#pragma unroll
  for (int j = 0; j < R; ++j) b.v[j] = xor_groups<M>(b.v[j]);
  b.d = xor_groups<M>(b.d);
  b.n = xor_groups<M>(b.n);
}

// Sweeps until one rotates nothing (or max_sweeps); nrot[2] are shared
// counters, zero on entry.  Every thread of the block calls this.
template <int R>
__device__ __forceinline__ void wl_sweeps(double* X, int k, double tol_rot, int max_sweeps, int* nrot) {
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, grp = lane >> 3, li = lane & 7;
  const int NB = 2 * ((k + 15) / 16);
  const bool active = wave < NB / 2;
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    int cnt = 0;
    for (int t = 0; t < NB - 1; ++t) {
      if (active) {
        const int P = circle_block(wave, t, NB), Q = circle_block(NB - 1 - wave, t, NB);
        WlCol<R> a, b;
        if (t == 0) {  // round 0 holds every block once: pairs inside the blocks
          wl_load<R>(X, k, P * 8 + grp, li, a);
          wl_intra_all<R>(a, grp, tol_rot, cnt);
          wl_store<R>(X, k, P * 8 + grp, li, a);
          wl_load<R>(X, k, Q * 8 + grp, li, b);
          wl_intra_all<R>(b, grp, tol_rot, cnt);
          wl_store<R>(X, k, Q * 8 + grp, li, b);
        }
        wl_load<R>(X, k, P * 8 + grp, li, a);  // the same lanes wrote these rows: in order
        wl_load<R>(X, k, Q * 8 + grp, li, b);
        // group g pairs its P column with the Q column of group g ^ gray(u)
        wl_cross<R>(a, b, tol_rot, cnt);
        wl_shift<R, 1>(b);
        wl_cross<R>(a, b, tol_rot, cnt);
        wl_shift<R, 2>(b);
        wl_cross<R>(a, b, tol_rot, cnt);
        wl_shift<R, 1>(b);
        wl_cross<R>(a, b, tol_rot, cnt);
        wl_shift<R, 4>(b);
        wl_cross<R>(a, b, tol_rot, cnt);
        wl_shift<R, 1>(b);
        wl_cross<R>(a, b, tol_rot, cnt);
        wl_shift<R, 2>(b);
        wl_cross<R>(a, b, tol_rot, cnt);
        wl_shift<R, 1>(b);
        wl_cross<R>(a, b, tol_rot, cnt);
        wl_store<R>(X, k, P * 8 + grp, li, a);
        wl_store<R>(X, k, Q * 8 + (grp ^ 4), li, b);  // gray(7) = 4
      }
      __syncthreads();
    }
    if (li == 0 && cnt) atomicAdd(&nrot[sweep & 1], cnt);
    __syncthreads();
    const int n = nrot[sweep & 1];
    if (tid == 0) nrot[(sweep + 1) & 1] = 0;
    __syncthreads();
    if (n == 0) break;
  }
}

// Pivoted Cholesky of G (in X, k x k column-major, k <= KM <= 64) by one
// wave: lane i keeps row i of the trailing matrix in registers, so a step
// takes no work-group barrier.  The trailing matrix stays symmetric to the
// bit (both halves take the same products), so the pivot column is lane p's
// row, which lane p writes to LDS (every index static: an index chosen at
// run time would move the row to scratch).  Pivot rule, keys and arithmetic
// are those of the register-tile form (the largest remaining diagonal,
// smallest index on ties; l = G(:, p) / sqrt(G(p, p)); every update the same
// fused multiply-add), so L is the same to the bit; column j of X receives l
// in the original row order.  Returns the rank.
template <int KM>
__device__ __forceinline__ int chol_wave(double* X, int k) {
  __shared__ double lb[64];  // the step's column l, one entry per lane
  __shared__ double cp[64];  // the pivot's row of the trailing matrix
  const int lane = threadIdx.x & 63;
  const bool row = lane < k;
  double g[KM];
#pragma unroll
  for (int c = 0; c < KM; ++c) g[c] = (row && c < k) ? X[c * k + lane] : 0.0;
  double dg = row ? X[lane * k + lane] : 0.0;
  bool used = !row;
  int rk = k;
  for (int j = 0; j < k; ++j) {
    long long key = used ? LLONG_MIN : ((__double_as_longlong(dg) & ~0x3FFLL) | (1023 - lane));
#pragma unroll
    for (int st = 0; st < 4; ++st) {
      const long long o = __double_as_longlong(dpp_bits(__longlong_as_double(key), st));
      key = o > key ? o : key;
    }
    long long o = __double_as_longlong(swizzle64<0x401F>(__longlong_as_double(key)));
    key = o > key ? o : key;
    o = __double_as_longlong(xor32(__longlong_as_double(key)));
    key = o > key ? o : key;
    const int p = __builtin_amdgcn_readfirstlane(key == LLONG_MIN ? -1 : 1023 - int(key & 0x3FF));
    double pv = 0.0;
    if (p >= 0) {
      const long long b = __double_as_longlong(dg);
      const int lo = __builtin_amdgcn_readlane(int(b & 0xffffffffLL), p);
      const int hi = __builtin_amdgcn_readlane(int(b >> 32), p);
      pv = __longlong_as_double((long long)(unsigned)lo | ((long long)hi << 32));
    }
    if (!(pv > 0.0)) {  // uniform: the rest of the matrix is zero or negative noise
      rk = j;
      break;
    }
    const double isq = fast_rsqrt(pv), sq = pv * isq;
    if (lane == p) {
#pragma unroll
      for (int c = 0; c < KM; ++c) cp[c] = g[c];
    }
    __builtin_amdgcn_wave_barrier();  // one wave: its LDS accesses complete in order
    const double gp = cp[lane];
    const double l = lane == p ? sq : (used ? 0.0 : gp * isq);
    if (lane == p) {
      used = true;
      dg = -1.0;
    } else {
      dg -= l * l;  // pivoted rows hold -1 and l = 0
    }
    if (row) X[j * k + lane] = l;
    lb[lane] = l;
    __builtin_amdgcn_wave_barrier();  // one wave: its LDS accesses complete in order
#pragma unroll
    for (int c = 0; c < KM; ++c) g[c] -= l * lb[c];
    __builtin_amdgcn_wave_barrier();  // every read of lb precedes the next step's writes
  }
  return rk;
}

__global__ __launch_bounds__(kGramEigThreads) void k_gram_eig(double* __restrict__ base,
                                                            const SmallJacobi* __restrict__ js,
                                                            int* __restrict__ flags, double tol, int max_sweeps,
                                                            int precond) {
  extern __shared__ double lds[];
  __shared__ int nrot[2];
  __shared__ int nviol, nabs, best;
  const SmallJacobi J = js[blockIdx.x];
  const int k = J.k;
  const int ke = k + (k & 1);
  const int tid = threadIdx.x;
  double* X = lds;                         // k x ke, column-major
  double* nrm = lds + int64_t(k) * ke;     // ke
  double* vec = nrm + ke;                  // ke
  double* hh = vec + ke;                   // ke
  int* perm = reinterpret_cast<int*>(hh + ke);  // ke
  double* Gg = base + J.y_off;
  double* Wg = base + J.w_off;
  const int npairs = ke / 2;
  const int gsz = gram_eig_gsz(k, int(blockDim.x));
  if (k <= 0) return;
  if (k > kGramEigMaxK) {  // never launched so by the host
    for (int64_t e = tid; e < int64_t(k) * k; e += blockDim.x)
      Wg[(e / k) * J.ldw + e % k] = (e / k == e % k) ? 1.0 : 0.0;
    if (tid == 0 && J.flag_idx >= 0) flags[J.flag_idx] = -1;
    return;
  }
  for (int e = tid; e < k * ke; e += blockDim.x) {
    const int r = e % k, c = e / k;
    X[e] = c < k ? Gg[int64_t(c) * J.ldy + r] : 0.0;
  }
  for (int e = tid; e < ke; e += blockDim.x) perm[e] = e;
  if (tid == 0) {
    nviol = 0;
    nabs = 0;
    nrot[0] = nrot[1] = 0;
  }
  __syncthreads();
  // ---- 1. orthogonality test on G -------------------------------------------
  {
    double dm = 0.0;
    for (int i = tid; i < k; i += blockDim.x) dm = fmax(dm, X[i * k + i]);
    if (tid == 0) nrm[0] = 0.0;  // scratch for the maximum (nrm is free until the Jacobi ends)
    __syncthreads();
    for (int off = 32; off > 0; off >>= 1) dm = fmax(dm, __shfl_xor(dm, off));
    if ((tid & 63) == 0) atomicMax(reinterpret_cast<unsigned long long*>(nrm),
                                   static_cast<unsigned long long>(__double_as_longlong(dm)));
    __syncthreads();
    // the two tests on squares: g^2 > tol^2 g_ii g_jj (relative) and
    // g^2 > tol^2 max(g_ii, g_jj) max_l g_ll (absolute); wave w scans the
    // upper part of columns w, w + nwave, ...
    const double g2max = nrm[0], t2 = tol * tol;
    const int lane = tid & 63, nwave = blockDim.x >> 6;
    int my = 0, mya = 0;
    for (int c = tid >> 6; c < k; c += nwave) {
      const double gjj = X[c * k + c];
      for (int r = lane; r < c; r += 64) {
        const double gii = X[r * k + r], g = X[c * k + r], g2 = g * g;
        if (gii > 0.0 && gjj > 0.0 && g2 > t2 * gii * gjj) {
          if (g2 > t2 * fmax(gii, gjj) * g2max)
            ++my;
          else
            ++mya;
        }
      }
    }
    if (my) atomicAdd(&nviol, my);
    if (mya) atomicAdd(&nabs, mya);
  }
  __syncthreads();
  const bool ident = nviol == 0 || k < 2;
  if (!ident && precond) {
    // ---- 2. pivoted Cholesky, G held in registers ------------------------------
    // Up to k = 64 one wave does it (chol_wave).  Wider, thread t owns an 8 x 5
    // tile of G.  Step j picks the largest remaining diagonal p (smallest
    // index on ties), the owners of column p publish l = G(:, p) / sqrt(G(p,
    // p)) (zero on rows already pivoted), every tile takes the rank-one update
    // and vec keeps the remaining diagonal.  Column j of X receives l in the
    // original row order, so X X^T = G and the Jacobi below needs no
    // permutation; perm flags the pivoted rows here and is reset to the
    // identity afterwards.
    __shared__ int rk_sh;
    int rk = k;
    if (k <= 64) {
      if (tid < 64) {
        const int r = k <= 32 ? chol_wave<32>(X, k) : chol_wave<64>(X, k);
        if (tid == 0) rk_sh = r;
      }
      __syncthreads();
      rk = rk_sh;
    } else {
      constexpr int TR = 8, TC = 5;
      const int ntc = (k + TC - 1) / TC;
      const bool owner = tid < ((k + TR - 1) / TR) * ntc;
      const int tr = owner ? tid / ntc : 0, tc = owner ? tid % ntc : 0;
      double g[TR][TC];
#pragma unroll
      for (int a = 0; a < TR; ++a)
#pragma unroll
        for (int b = 0; b < TC; ++b) {
          const int i = tr * TR + a, l = tc * TC + b;
          g[a][b] = (owner && i < k && l < k) ? X[l * k + i] : 0.0;
        }
      // Pivot keys: the top 54 bits of the remaining diagonal and the inverted
      // index in the low 10 bits (k <= 1023), so the largest key is the largest
      // value, smallest index on ties; pivoted entries hold -1, and non-positive
      // values never win over positive ones under a signed compare.  Thread i <
      // k keeps row i; each wave holding rows reduces its keys and publishes the
      // maximum in wkey[parity of the step it serves], which every thread reads
      // after the step's second barrier.
      __shared__ long long wkey[2][kGramEigThreads / 64];
      const int nwv = (k + 63) >> 6;  // waves that hold rows
      auto publish = [&](int slot) {
        long long key = LLONG_MIN;
        if (tid < k) key = (__double_as_longlong(vec[tid]) & ~0x3FFLL) | (1023 - tid);
#pragma unroll
        for (int st = 0; st < 4; ++st) {
          const long long o = __double_as_longlong(dpp_bits(__longlong_as_double(key), st));
          key = o > key ? o : key;
        }
        long long o = __double_as_longlong(swizzle64<0x401F>(__longlong_as_double(key)));
        key = o > key ? o : key;
        o = __double_as_longlong(xor32(__longlong_as_double(key)));
        key = o > key ? o : key;
        if ((tid & 63) == 0) wkey[slot][tid >> 6] = key;
      };
      if (tid < k) {
        vec[tid] = X[tid * k + tid];
        perm[tid] = 0;
      }
      if ((tid >> 6) < nwv) publish(0);
      __syncthreads();  // every read of G from X precedes the L columns written below
      // Each step: (1) every thread takes the pivot p from the wave maxima and
      // the owners of column p copy it to hh; (2) thread i forms l_i, writes it
      // to nrm and to column j of X, updates vec[i] and the waves publish the
      // next keys; (3) the tiles take the rank-one update.  Two barriers a step:
      // hh and nrm are written one phase after their last reads, behind a
      // barrier, and the wave maxima alternate between two slots.
      for (int j = 0; j < k; ++j) {
        long long top = wkey[j & 1][0];
        for (int w = 1; w < nwv; ++w) top = wkey[j & 1][w] > top ? wkey[j & 1][w] : top;
        const int p = top == LLONG_MIN ? -1 : 1023 - int(top & 0x3FF);
        const double pmx = p >= 0 ? vec[p] : 0.0;
        if (!(pmx > 0.0)) {  // uniform: the rest of the matrix is zero or negative noise
          rk = j;
          break;
        }
        const double isq = fast_rsqrt(pmx), sq = pmx * isq;
        if (owner && tc == p / TC) {
          // column p - tc TC of the tile by 0/1 weights: a select on the index
          // would be folded into an indexed load and move the tile to scratch
          double m[TC];
#pragma unroll
          for (int b = 0; b < TC; ++b) m[b] = (p - tc * TC == b) ? 1.0 : 0.0;
#pragma unroll
          for (int a = 0; a < TR; ++a) {
            double v = 0.0;
#pragma unroll
            for (int b = 0; b < TC; ++b) v += g[a][b] * m[b];
            if (tr * TR + a < k) hh[tr * TR + a] = v;
          }
        }
        __syncthreads();
        if (tid < k) {
          const int i = tid;
          const double l = i == p ? sq : (perm[i] ? 0.0 : hh[i] * isq);
          nrm[i] = l;
          X[j * k + i] = l;
          if (i == p) {
            vec[p] = -1.0;
            perm[p] = 1;
          } else {
            vec[i] -= l * l;  // pivoted rows hold -1 and l = 0
          }
        }
        if ((tid >> 6) < nwv) publish((j + 1) & 1);
        __syncthreads();
        if (owner) {
          double li[TR], lc[TC];
#pragma unroll
          for (int a = 0; a < TR; ++a) li[a] = tr * TR + a < k ? nrm[tr * TR + a] : 0.0;
#pragma unroll
          for (int b = 0; b < TC; ++b) lc[b] = tc * TC + b < k ? nrm[tc * TC + b] : 0.0;
#pragma unroll
          for (int a = 0; a < TR; ++a)
#pragma unroll
            for (int b = 0; b < TC; ++b) g[a][b] -= li[a] * lc[b];
        }
      }
    }
    // columns from the rank on are zero; perm back to the identity
    for (int e = rk * k + tid; e < k * k; e += blockDim.x) X[e] = 0.0;
    for (int i = tid; i < ke; i += blockDim.x) perm[i] = i;
    __syncthreads();
  }
  // ---- 3. one-sided Jacobi on the columns of X ------------------------------
  const double tol_rot = tol * (1.0 / 16.0);
  const int pi = tid / gsz, li = tid % gsz;
  if (!ident) {
    if (k <= 16)
      wl_sweeps<2>(X, k, tol_rot, max_sweeps, nrot);
    else if (k <= 32)
      wl_sweeps<4>(X, k, tol_rot, max_sweeps, nrot);
    else if (k <= 48)
      wl_sweeps<6>(X, k, tol_rot, max_sweeps, nrot);
    else if (k <= 64)
      wl_sweeps<8>(X, k, tol_rot, max_sweeps, nrot);
    else if (k <= 80)
      wl_sweeps<10>(X, k, tol_rot, max_sweeps, nrot);
    else if (k <= 96)
      wl_sweeps<12>(X, k, tol_rot, max_sweeps, nrot);
    else if (k <= 112)
      wl_sweeps<14>(X, k, tol_rot, max_sweeps, nrot);
    else if (k <= 128)
      wl_sweeps<16>(X, k, tol_rot, max_sweeps, nrot);
    else
      wl_sweeps<18>(X, k, tol_rot, max_sweeps, nrot);
  }
  if (!ident) {
    // ---- 4. normalise, complete zero columns --------------------------------
    // column norms: the pairs of round 0 cover every column once
    if (pi < npairs) {
      const int a = circle_block(pi, 0, ke), b = circle_block(ke - 1 - pi, 0, ke);
      double al = 0.0, be = 0.0;
      double zero = 0.0;
      for (int l = li; l < k; l += gsz) {
        al += X[a * k + l] * X[a * k + l];
        be += X[b * k + l] * X[b * k + l];
      }
      group_sum3(al, be, zero, gsz);
      if (li == 0) {
        nrm[a] = sqrt(al);
        nrm[b] = sqrt(be);
      }
    }
    __syncthreads();
    for (int e = tid; e < k * k; e += blockDim.x) {
      const double v = nrm[e / k];
      X[e] = v > 1e-300 ? X[e] / v : 0.0;
    }
    __syncthreads();
    for (int c = 0; c < k; ++c) {
      if (nrm[c] > 1e-300) continue;  // uniform: nrm is shared
      for (int r = tid; r < k; r += blockDim.x) {
        double q = 1.0;
        for (int j = 0; j < k; ++j) q -= X[j * k + r] * X[j * k + r];
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
        for (int j = 0; j < k; ++j) q -= X[j * k + r] * X[j * k + e0];
        vec[r] = q;
      }
      __syncthreads();
      for (int j = tid; j < k; j += blockDim.x) {
        double q = 0.0;
        for (int r = 0; r < k; ++r) q += X[j * k + r] * vec[r];
        hh[j] = q;
      }
      __syncthreads();
      for (int r = tid; r < k; r += blockDim.x) {
        double q = vec[r];
        for (int j = 0; j < k; ++j) q -= X[j * k + r] * hh[j];
        vec[r] = q;
      }
      __syncthreads();
      if (tid == 0) {
        double q = 0.0;
        for (int r = 0; r < k; ++r) q += vec[r] * vec[r];
        nrm[c] = sqrt(q);
      }
      __syncthreads();
      for (int r = tid; r < k; r += blockDim.x) X[c * k + r] = vec[r] / nrm[c];
      __syncthreads();
    }
  }
  // ---- 5. W = P Q ---------------------------------------------------------------
  for (int e = tid; e < k * k; e += blockDim.x) {
    const int r = e % k, c = e / k;
    if (ident)
      Wg[int64_t(c) * J.ldw + r] = (r == c) ? 1.0 : 0.0;
    else
      Wg[int64_t(c) * J.ldw + perm[r]] = X[e];
  }
  if (tid == 0 && J.flag_idx >= 0) flags[J.flag_idx] = nviol;
  if (tid == 0 && J.aux_idx >= 0) flags[J.aux_idx] = nabs;
}

// First-order Gram-Schmidt step on columns whose Gram matrix is G (k x k,
// leading dimension ldy): T = I - M with M_ij = g_ij / g_ii when column i is
// longer than column j (ties: lower index), so Z T removes from every column
// its components along the longer ones.  T is written k x k at w_off
// (leading dimension ldw).  One work-group per matrix.
__global__ __launch_bounds__(256) void k_gs_step(double* __restrict__ base, const SmallJacobi* __restrict__ js) {
  const SmallJacobi J = js[blockIdx.x];
  const int k = J.k;
  const int j0 = blockIdx.y * 8;  // eight columns of T per work-group
  if (j0 >= k) return;
  const double* G = base + J.y_off;
  double* T = base + J.w_off;
  double dj[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) dj[c] = j0 + c < k ? G[int64_t(j0 + c) * J.ldy + j0 + c] : 0.0;
  for (int i = threadIdx.x; i < k; i += blockDim.x) {
    const double di = G[int64_t(i) * J.ldy + i];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int j = j0 + c;
      if (j < k) {
        double v = i == j ? 1.0 : 0.0;
        if (i != j && di > 0.0 && (di > dj[c] || (di == dj[c] && i < j))) v = -G[int64_t(j) * J.ldy + i] / di;
        T[int64_t(j) * J.ldw + i] = v;
      }
    }
  }
}

void gs_step(hipStream_t st, double* base, const void* descs, int n, int kmax) {
  if (n <= 0 || kmax <= 0) return;
  hipLaunchKernelGGL(k_gs_step, dim3(n, (kmax + 7) / 8), dim3(256), 0, st, base,
                     static_cast<const SmallJacobi*>(descs));
}

void gram_eig(hipStream_t st, double* base, const void* descs, int n, int* flags, double tol, int max_sweeps,
              int precond, size_t lds_bytes) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_gram_eig, dim3(n), dim3(kGramEigThreads), lds_bytes, st, base,
                     static_cast<const SmallJacobi*>(descs), flags, tol, max_sweeps, precond);
}

// X, nrm, vec, hh and perm, and 16 doubles beyond them: the wave-local loads
// of the last column read up to 8 R - k - 1 <= 15 rows past its end (masked).
size_t gram_eig_lds_bytes(int k) {
  const int64_t ke = k + (k & 1);
  return size_t(int64_t(k) * ke + 3 * ke + 16) * sizeof(double) + size_t(ke) * sizeof(int);
}

size_t jacobi_small_lds_limit() {
  static const size_t limit = [] {
    int dev = 0;
    int v = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) v = 65536;
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_jacobi_small),
                              hipFuncAttributeMaxDynamicSharedMemorySize, v - 64);
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_gram_eig),
                              hipFuncAttributeMaxDynamicSharedMemorySize, v - 64);
    return size_t(v);
  }();
  return limit;
}

// ---- finalize ----------------------------------------------------------------
// aux = base + w_off + kp*kp holds 2*kp + ell doubles: column norms, ranks and
// the completion vector.

// One work-group per sector: singular values (column norms of Y), their rank
// in descending order (ties broken by column index) and the sorted values.
// Column norms of Y: grid (sector, chunk of 16 columns); each wave sums
// four columns at once (four independent loads per row step).
__global__ __launch_bounds__(256) void k_svd_colnorms(double* __restrict__ base, const SvdSector* __restrict__ secs) {
  const SvdSector s = secs[blockIdx.x];
  const int k = s.k, ell = s.ell, kp = s.nb * kSvdBlock;
  const int c0 = blockIdx.y * 16 + (threadIdx.x >> 6) * 4;
  if (c0 >= k) return;
  const double* Y = base + s.y_off;
  double* sig = base + s.w_off + int64_t(kp) * kp;
  const int lane = threadIdx.x & 63;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  const double* col[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) col[c] = Y + int64_t(min(c0 + c, k - 1)) * ell;  // columns past k: masked below
  for (int l = lane; l < ell; l += 256) {  // 16 loads in flight per lane
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int r = l + 64 * u;
      const int rr = r < ell ? r : ell - 1;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const double x = col[c][rr];
        acc[c] += (r < ell && c0 + c < k) ? x * x : 0.0;
      }
    }
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    double a = acc[c];
    for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off);
    if (lane == 0 && c0 + c < k) sig[c0 + c] = sqrt(a);
  }
}

// Rank of every singular value of a sector in descending order (ties broken
// by column index) and the sorted values.  Work-group (sector, y) is one wave
// that ranks values 64 y ... 64 y + 63 against all k, compared from LDS 1024
// at a time.
constexpr int kNormsWave = 64;
__global__ __launch_bounds__(kNormsWave) void k_svd_norms(double* __restrict__ base,
                                                           const SvdSector* __restrict__ secs) {
  constexpr int kChunk = 1024;
  __shared__ double chunk[kChunk];
  const SvdSector s = secs[blockIdx.x];
  const int k = s.k, kp = s.nb * kSvdBlock;
  const int j0 = blockIdx.y * kNormsWave;
  if (j0 >= k) return;  // uniform
  double* aux = base + s.w_off + int64_t(kp) * kp;
  double* sig = aux;
  double* rnk = aux + kp;
  double* S = base + s.s_off;
  const int tid = threadIdx.x;
  const int j = j0 + tid;
  const double sj = j < k ? sig[j] : 0.0;
  int r = 0;
  for (int i0 = 0; i0 < k; i0 += kChunk) {
    const int n = min(kChunk, k - i0);
    __syncthreads();  // the previous chunk has been read
    for (int t = tid; t < n; t += blockDim.x) chunk[t] = sig[i0 + t];
    __syncthreads();
    for (int t = 0; t < n; ++t) {
      const double si = chunk[t];
      r += (si > sj) || (si == sj && i0 + t < j);
    }
  }
  if (j < k) {
    rnk[j] = double(r);
    if (s.eig_shift)
      S[k - 1 - r] = sj - base[s.shift_off];
    else
      S[r] = sj;
  }
}

// Scatter of one 64x64 tile of a column-major source (Y: ell x k, or W: k x k
// leading block of kp x kp) into the row-major output with permuted, scaled
// columns, transposed through LDS so both the reads and the writes are
// coalesced.  task = (sector, part, row0, col0); part 0: Y -> Z side
// (columns divided by their norm, zero when numerically null), part 1: W.
__global__ __launch_bounds__(256) void k_svd_write(double* __restrict__ base, const SvdSector* __restrict__ secs,
                                                   const int4* __restrict__ tasks) {
  __shared__ double tile[64][65];
  __shared__ int perm[64];
  __shared__ double scl[64];
  const int4 t = tasks[blockIdx.x];
  const SvdSector s = secs[t.x];
  const int k = s.k, kp = s.nb * kSvdBlock;
  const double* aux = base + s.w_off + int64_t(kp) * kp;
  const double* sig = aux;
  const double* rnk = aux + kp;
  const bool zpart = t.y == 0;
  const int rows = zpart ? s.ell : k;
  const int64_t ld = zpart ? s.ell : kp;
  const double* src = base + (zpart ? s.y_off : s.w_off);
  double* dst;
  if (s.eig_shift)
    dst = base + s.u_off;
  else
    dst = base + ((zpart != (s.transposed != 0)) ? s.u_off : s.v_off);
  const int r0 = t.z, c0 = t.w;
  const int tid = threadIdx.x;
  const double smax = base[s.s_off + (s.eig_shift ? k - 1 : 0)];
  if (tid < 64) {
    const int c = c0 + tid;
    if (c < k) {
      const int r = int(rnk[c]);
      perm[tid] = s.eig_shift ? k - 1 - r : r;
      if (zpart) {
        const double sj = sig[c];
        const double thr = (s.eig_shift ? 0.0 : smax) * double(s.ell) * DBL_EPSILON;
        scl[tid] = (sj > thr && sj > 0.0) ? 1.0 / sj : 0.0;
      } else {
        scl[tid] = 1.0;
      }
    }
  }
  for (int i = 0; i < 16; ++i) {
    const int cl = (tid >> 6) + 4 * i;
    const int rl = tid & 63;
    const int r = r0 + rl, c = c0 + cl;
    tile[cl][rl] = (r < rows && c < k) ? src[int64_t(c) * ld + r] : 0.0;
  }
  __syncthreads();
  for (int i = 0; i < 16; ++i) {
    const int rl = (tid >> 6) + 4 * i;
    const int cl = tid & 63;
    const int r = r0 + rl, c = c0 + cl;
    if (r < rows && c < k) dst[int64_t(r) * k + perm[cl]] = tile[cl][rl] * scl[cl];
  }
}

// Orthonormal completion of the Z-side columns whose singular value is
// numerically zero, columns t < limit: starting from unit vectors, classical
// Gram-Schmidt twice against all earlier columns (the dot products of one
// pass in one step, thread per column; the update thread per row), keeping
// the first candidate whose residual norm exceeds 1/2.  One work-group per
// sector.
__global__ __launch_bounds__(256) void k_svd_complete(double* __restrict__ base,
                                                      const CompleteTask* __restrict__ ts) {
  __shared__ double sh[256];
  __shared__ int first_sh;
  const CompleteTask T = ts[blockIdx.x];
  const int k = T.k, ell = T.ell, lim = T.limit < k ? T.limit : k;
  if (lim <= 0) return;
  const double* S = base + T.s_off;
  double* v = base + T.v_off;
  double* p = base + T.p_off;
  double* Z = base + T.z_off;  // ell x k row-major
  const double null_thr = S[0] * double(ell) * DBL_EPSILON;
  const int tid = threadIdx.x;
  // the first numerically zero value below the limit (lim when none)
  if (tid == 0) first_sh = lim;
  __syncthreads();
  for (int t = tid; t < lim; t += blockDim.x)
    if (!(S[t] > null_thr && S[t] > 0.0)) atomicMin(&first_sh, t);
  __syncthreads();
  const int first_null = first_sh;
  for (int t = first_null; t < lim; ++t) {
    for (int cand = 0; cand < ell; ++cand) {
      const int e0 = (t + cand) % ell;
      for (int l = tid; l < ell; l += blockDim.x) v[l] = (l == e0) ? 1.0 : 0.0;
      __syncthreads();
      for (int pass = 0; pass < 2; ++pass) {
        for (int c = tid; c < t; c += blockDim.x) {
          double acc = 0.0;
          for (int l = 0; l < ell; ++l) acc = fma(Z[int64_t(l) * k + c], v[l], acc);
          p[c] = acc;
        }
        __syncthreads();
        for (int l = tid; l < ell; l += blockDim.x) {
          const double* zr = Z + int64_t(l) * k;
          double acc = 0.0;
          for (int c = 0; c < t; ++c) acc = fma(zr[c], p[c], acc);
          v[l] -= acc;
        }
        __syncthreads();
      }
      double acc = 0.0;
      for (int l = tid; l < ell; l += blockDim.x) acc += v[l] * v[l];
      const double nrm = sqrt(wg_sum(acc, sh));
      if (nrm > 0.5) {
        for (int l = tid; l < ell; l += blockDim.x) Z[int64_t(l) * k + t] = v[l] / nrm;
        __syncthreads();
        break;
      }
    }
  }
}

void svd_complete(hipStream_t st, double* base, const CompleteTask* tasks, int n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_svd_complete, dim3(n), dim3(256), 0, st, base, tasks);
}

void svd_finalize(hipStream_t st, double* base, const SvdSector* secs, const SvdSector* host_secs, int nsec,
                  const int4* tasks, int ntasks) {
  if (nsec <= 0) return;
  int kmax = 1;
  for (int i = 0; i < nsec; ++i) kmax = std::max(kmax, host_secs[i].k);
  hipLaunchKernelGGL(k_svd_colnorms, dim3(nsec, (kmax + 15) / 16), dim3(256), 0, st, base, secs);
  hipLaunchKernelGGL(k_svd_norms, dim3(nsec, (kmax + kNormsWave - 1) / kNormsWave), dim3(kNormsWave), 0, st, base, secs);
  if (ntasks > 0) hipLaunchKernelGGL(k_svd_write, dim3(ntasks), dim3(256), 0, st, base, secs, tasks);
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
// The values of all sectors are gathered, sorted descending (bitonic, in LDS
// when 2^ceil(log2 total) doubles fit, else in the scratch area) and the cut
// is found by the reference's sequential tail accumulation.
template <bool IN_LDS>
__global__ __launch_bounds__(1024) void k_truncate(double* __restrict__ base, const int64_t* __restrict__ s_offs,
                                                   const int32_t* __restrict__ ks, const int64_t* __restrict__ s_pos,
                                                   int nsec, int64_t scratch_off,
                                                   int64_t total, double tol, double pw, int64_t min_size,
                                                   int64_t max_size, int32_t* __restrict__ kept,
                                                   double* __restrict__ thr_out) {
  extern __shared__ double vl[];
  __shared__ double thr_sh;
  double* v = IN_LDS ? vl : base + scratch_off;
  int64_t n2 = 1;
  while (n2 < total) n2 <<= 1;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6, nw = blockDim.x >> 6;
  // gather (sector lists concatenated in order, s_pos[i] = start of sector i),
  // one wave per sector, pad with -inf
  for (int i = wave; i < nsec; i += nw) {
    const double* src = base + s_offs[i];
    const int64_t pos = s_pos[i];
    const int n = ks[i];
    for (int j = lane; j < n; j += 64) v[pos + j] = src[j];
  }
  for (int64_t j = total + tid; j < n2; j += blockDim.x) v[j] = -INFINITY;
  __syncthreads();
  // bitonic sort, descending (n2 < 2^31; strides are powers of two, so the
  // pair index is formed with masks, not divisions)
  const int m2 = int(n2);
  for (int size = 2; size <= m2; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int i = tid; i < m2 / 2; i += blockDim.x) {
        const int lo = ((i & ~(stride - 1)) << 1) | (i & (stride - 1));
        const int hi = lo + stride;
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
    // the same steps in the same order, the next eight values read ahead
    // (a dependent read per step would dominate the scan); the square and
    // the general power are separate loops, so that the common pw = 2 scan
    // evaluates no pow()
    auto scan = [&](auto power) {
      const double toln = power(tol);
      int64_t last = total - 1;
      double trunc = power(fabs(v[last]));
      for (bool done = false; !done;) {
        double buf[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) buf[j] = last - 1 - j >= 0 ? v[last - 1 - j] : 0.0;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          if (!(last >= min_size && last > 0) || (trunc > toln && last < max_size)) {
            done = true;
            break;
          }
          --last;
          trunc += power(fabs(buf[j]));
        }
      }
      return last;
    };
    const int64_t last = pw == 2.0 ? scan([](double x) { return x * x; })
                                   : scan([pw](double x) { return pow(x, pw); });
    const double sl = v[last];
    thr_sh = sl - 2.0 * sl * DBL_EPSILON;
    thr_out[0] = thr_sh;
  }
  __syncthreads();
  const double thr = thr_sh;
  // kept[i] = index of the first value of sector i that is not above thr
  // (its length when none), one wave per sector, 64 values per ballot; the
  // chunks are all read (no early exit) so that their loads overlap
  for (int i = wave; i < nsec; i += nw) {
    const double* src = base + s_offs[i];
    const int n = ks[i];
    int c = n;
#pragma unroll 4
    for (int j0 = 0; j0 < n; j0 += 64) {
      const int j = j0 + lane;
      const unsigned long long m = __ballot(j < n && !(src[j < n ? j : 0] > thr));
      if (m && c == n) c = j0 + __ffsll((long long)m) - 1;
    }
    if (lane == 0) kept[i] = c;
  }
}

void truncate_select(hipStream_t st, double* base, const int64_t* s_offs, const int32_t* ks, const int64_t* s_pos,
                     int nsec, int64_t scratch_off, int64_t total, double tol, double pw, int64_t min_size,
                     int64_t max_size, int32_t* kept, double* thr_out) {
  int64_t n2 = 1;
  while (n2 < total) n2 <<= 1;
  if (n2 > (int64_t(1) << 30)) throw std::runtime_error("truncation: more than 2^30 values");
  const size_t lds = size_t(n2) * sizeof(double);
  static const size_t limit = [] {
    int dev = 0, v = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) v = 65536;
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_truncate<true>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, v - 256);
    return size_t(v) - 256;
  }();
  if (lds <= limit)
    hipLaunchKernelGGL(k_truncate<true>, dim3(1), dim3(1024), lds, st, base, s_offs, ks, s_pos, nsec, scratch_off,
                       total, tol, pw, min_size, max_size, kept, thr_out);
  else
    hipLaunchKernelGGL(k_truncate<false>, dim3(1), dim3(1024), 0, st, base, s_offs, ks, s_pos, nsec, scratch_off,
                       total, tol, pw, min_size, max_size, kept, thr_out);
}

}  // namespace kern
}  // namespace qtt
