// Grouped block-sparse GEMM for gfx950 (CDNA4).
//
// One work-group computes one 64x64 tile of one output block and walks that
// block's list of (A_p, B_p) pairs in order, so every output block is written
// exactly once with a fixed accumulation order (no atomics: results do not
// depend on what else runs on the device).  The permute gather of the
// reference's permute_bl (sources/btensor.cpp:1843-1894) is folded into the
// tile loads through per-pair (row, column) element strides.
//
// FP64 path: v_mfma_f64_16x16x4_f64.  Each of the 4 waves owns a 32x32
// quadrant = 2x2 MFMA tiles; LDS holds a 64x16 A tile and a 16x64 B tile
// (k-major, row pitch 80 doubles so the two 16-lane halves of a 32-lane
// LDS read phase fall in opposite bank halves).
#include <map>
#include <mutex>

#include "../kernels.hpp"

namespace qtt {
namespace kern {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int BM = 64, BN = 64, BK = 16;
constexpr int LDP = BM + 16;  // LDS row pitch (doubles)

// Loads the A tile (BM x BK, element (m,k) at A[m*rs + k*cs]) into As[k][m]
// and the B tile (BK x BN, element (k,n) at B[k*rs + n*cs]) into Bs[k][n].
template <class T>
__device__ __forceinline__ void load_tiles(const double* __restrict__ A, const double* __restrict__ B,
                                           const GemmPair& pr, int M, int N, int m0, int n0, int k0, int tid,
                                           T (*As)[LDP], T (*Bs)[LDP], T (*As2)[LDP], T (*Bs2)[LDP]) {
  const int K = pr.K;
  if (pr.a_cs == 1) {  // row-major A: 4 consecutive k per thread
    const int m = tid >> 2, kq = (tid & 3) * 4;
    const int gm = m0 + m;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int gk = k0 + kq + j;
      double v = (gm < M && gk < K) ? A[int64_t(gm) * pr.a_rs + gk] : 0.0;
      if constexpr (sizeof(T) == 8) {
        As[kq + j][m] = v;
      } else {
        float hi = float(v);
        As[kq + j][m] = hi;
        As2[kq + j][m] = float(v - double(hi));
      }
    }
  } else {  // column-major or general strides: 4 consecutive m per thread
    const int k = tid >> 4, mq = (tid & 15) * 4;
    const int gk = k0 + k;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int gm = m0 + mq + j;
      double v = (gm < M && gk < K) ? A[int64_t(gm) * pr.a_rs + int64_t(gk) * pr.a_cs] : 0.0;
      if constexpr (sizeof(T) == 8) {
        As[k][mq + j] = v;
      } else {
        float hi = float(v);
        As[k][mq + j] = hi;
        As2[k][mq + j] = float(v - double(hi));
      }
    }
  }
  if (pr.b_cs == 1) {  // row-major B: 4 consecutive n per thread
    const int k = tid >> 4, nq = (tid & 15) * 4;
    const int gk = k0 + k;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int gn = n0 + nq + j;
      double v = (gn < N && gk < K) ? B[int64_t(gk) * pr.b_rs + gn] : 0.0;
      if constexpr (sizeof(T) == 8) {
        Bs[k][nq + j] = v;
      } else {
        float hi = float(v);
        Bs[k][nq + j] = hi;
        Bs2[k][nq + j] = float(v - double(hi));
      }
    }
  } else {  // column-major or general: 4 consecutive k per thread
    const int n = tid >> 2, kq = (tid & 3) * 4;
    const int gn = n0 + n;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int gk = k0 + kq + j;
      double v = (gn < N && gk < K) ? B[int64_t(gk) * pr.b_rs + int64_t(gn) * pr.b_cs] : 0.0;
      if constexpr (sizeof(T) == 8) {
        Bs[kq + j][n] = v;
      } else {
        float hi = float(v);
        Bs[kq + j][n] = hi;
        Bs2[kq + j][n] = float(v - double(hi));
      }
    }
  }
}

// Register-staged variant of load_tiles for the FP64 kernel: fetch() issues
// the global loads of one (A, B) k-tile into registers, stash() writes them to
// LDS, so the loads of tile t+1 overlap the MFMAs of tile t.
struct Regs {
  double a[4], b[4];
};

__device__ __forceinline__ void fetch(const double* __restrict__ base, const GemmPair& pr, int M, int N, int m0,
                                      int n0, int k0, int tid, Regs& r) {
  const double* A = base + pr.a_off;
  const double* B = base + pr.b_off;
  const int K = pr.K;
  if (pr.a_cs == 1) {
    const int m = tid >> 2, kq = (tid & 3) * 4, gm = m0 + m;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int gk = k0 + kq + j;
      r.a[j] = (gm < M && gk < K) ? A[int64_t(gm) * pr.a_rs + gk] : 0.0;
    }
  } else {
    const int k = tid >> 4, mq = (tid & 15) * 4, gk = k0 + k;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int gm = m0 + mq + j;
      r.a[j] = (gm < M && gk < K) ? A[int64_t(gm) * pr.a_rs + int64_t(gk) * pr.a_cs] : 0.0;
    }
  }
  if (pr.b_cs == 1) {
    const int k = tid >> 4, nq = (tid & 15) * 4, gk = k0 + k;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int gn = n0 + nq + j;
      r.b[j] = (gn < N && gk < K) ? B[int64_t(gk) * pr.b_rs + gn] : 0.0;
    }
  } else {
    const int n = tid >> 2, kq = (tid & 3) * 4, gn = n0 + n;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int gk = k0 + kq + j;
      r.b[j] = (gn < N && gk < K) ? B[int64_t(gk) * pr.b_rs + int64_t(gn) * pr.b_cs] : 0.0;
    }
  }
}

__device__ __forceinline__ void stash(const GemmPair& pr, int tid, const Regs& r, double (*As)[LDP],
                                      double (*Bs)[LDP]) {
  if (pr.a_cs == 1) {
    const int m = tid >> 2, kq = (tid & 3) * 4;
#pragma unroll
    for (int j = 0; j < 4; ++j) As[kq + j][m] = r.a[j];
  } else {
    const int k = tid >> 4, mq = (tid & 15) * 4;
#pragma unroll
    for (int j = 0; j < 4; ++j) As[k][mq + j] = r.a[j];
  }
  if (pr.b_cs == 1) {
    const int k = tid >> 4, nq = (tid & 15) * 4;
#pragma unroll
    for (int j = 0; j < 4; ++j) Bs[k][nq + j] = r.b[j];
  } else {
    const int n = tid >> 2, kq = (tid & 3) * 4;
#pragma unroll
    for (int j = 0; j < 4; ++j) Bs[kq + j][n] = r.b[j];
  }
}

// Output element (row, col) of 16x16 MFMA tile register r in lane l:
//   v_mfma_f64_16x16x4_f64: row = 4*r + l/16, col = l%16
//   v_mfma_f32_16x16x4_f32: row = 4*(l/16) + r, col = l%16
// (measured on gfx950; the two layouts differ).
template <bool F64Layout>
__device__ __forceinline__ void store_tile(double* __restrict__ C, const GemmOut& o, int m0, int n0, int wm, int wn,
                                           int lane, const v4d (&acc)[2][2], double alpha, double beta) {
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm + i * 16 + (F64Layout ? 4 * r + (lane >> 4) : 4 * (lane >> 4) + r);
        const int col = n0 + wn + j * 16 + (lane & 15);
        if (row < o.M && col < o.N) {
          double* p = C + int64_t(row) * o.ldc + col;
          double v = alpha * acc[i][j][r];
          if (beta != 0.0) v += beta * (*p);
          *p = v;
        }
      }
}

__global__ __launch_bounds__(256) void k_gemm_f64(double* __restrict__ base, const GemmTile* __restrict__ tiles,
                                                  const GemmOut* __restrict__ outs,
                                                  const GemmPair* __restrict__ pairs, double alpha, double beta) {
  __shared__ double As[BK][LDP];
  __shared__ double Bs[BK][LDP];
  const GemmTile t = tiles[blockIdx.x];
  const GemmOut o = outs[t.out];
  const int m0 = t.tm * BM, n0 = t.tn * BN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  v4d acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};

  // flattened (pair, k-tile) iteration with one tile of register prefetch
  int p = o.pair_begin, k0 = 0;
  while (p < o.pair_end && pairs[p].K <= 0) ++p;
  if (p < o.pair_end) {
    GemmPair pr = pairs[p];
    Regs cur;
    fetch(base, pr, o.M, o.N, m0, n0, 0, tid, cur);
    while (true) {
      stash(pr, tid, cur, As, Bs);
      __syncthreads();
      // advance to the next tile and prefetch it
      int np = p, nk = k0 + BK;
      GemmPair npr = pr;
      if (nk >= pr.K) {
        nk = 0;
        ++np;
        while (np < o.pair_end && pairs[np].K <= 0) ++np;
        if (np < o.pair_end) npr = pairs[np];
      }
      const bool more = np < o.pair_end;
      Regs nxt;
      if (more) fetch(base, npr, o.M, o.N, m0, n0, nk, tid, nxt);
#pragma unroll
      for (int kk = 0; kk < BK; kk += 4) {
        const int kr = kk + (lane >> 4), c = lane & 15;
        const double a0 = As[kr][wm + c], a1 = As[kr][wm + 16 + c];
        const double b0 = Bs[kr][wn + c], b1 = Bs[kr][wn + 16 + c];
        acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
      }
      __syncthreads();
      if (!more) break;
      p = np;
      k0 = nk;
      pr = npr;
      cur = nxt;
    }
  }
  store_tile<true>(base + o.c_off, o, m0, n0, wm, wn, lane, acc, alpha, beta);
}

// ---- tile variants of the FP64 grouped GEMM ---------------------------------
// TM x TN output tile per work-group (4 waves in 2 x 2, each (TM/2) x (TN/2)
// as (TM/32) x (TN/32) MFMA 16x16 tiles), k in steps of KD (16 or 32)
// through a double-buffered LDS: the global loads of step t+1 are in
// registers while the MFMAs of step t run, then written to the other buffer,
// one barrier per step.  Every output element accumulates over k in the same
// order as k_gemm_f64 (4-wide MFMA chunks from each pair's k = 0, pairs in
// list order; the zero padding of a pair's last step adds exact zeros), so
// all variants give the same bits.
template <int TM, int TN, int KD>
struct TileRegs {
  double a[TM * KD / 256], b[TN * KD / 256];
};

template <int TM, int TN, int KD>
__device__ __forceinline__ void fetch_t(const double* __restrict__ base, const GemmPair& pr, int M, int N, int m0,
                                        int n0, int k0, int tid, TileRegs<TM, TN, KD>& r) {
  constexpr int EA = TM * KD / 256, EB = TN * KD / 256;  // elements per thread
  const double* A = base + pr.a_off;
  const double* B = base + pr.b_off;
  const int K = pr.K;
  if (pr.a_cs == 1) {  // k contiguous: KD/EA threads per row, EA consecutive k each
    constexpr int TPR = KD / EA;
    const int m = tid / TPR, kq = (tid % TPR) * EA, gm = m0 + m;
#pragma unroll
    for (int j = 0; j < EA; ++j) {
      const int gk = k0 + kq + j;
      r.a[j] = (gm < M && gk < K) ? A[int64_t(gm) * pr.a_rs + gk] : 0.0;
    }
  } else {  // m contiguous (or general): TM/EA threads per k, EA consecutive m each
    constexpr int TPK = TM / EA;
    const int k = tid / TPK, mq = (tid % TPK) * EA, gk = k0 + k;
#pragma unroll
    for (int j = 0; j < EA; ++j) {
      const int gm = m0 + mq + j;
      r.a[j] = (gm < M && gk < K) ? A[int64_t(gm) * pr.a_rs + int64_t(gk) * pr.a_cs] : 0.0;
    }
  }
  if (pr.b_cs == 1) {  // n contiguous: TN/EB threads per k, EB consecutive n each
    constexpr int TPK = TN / EB;
    const int k = tid / TPK, nq = (tid % TPK) * EB, gk = k0 + k;
#pragma unroll
    for (int j = 0; j < EB; ++j) {
      const int gn = n0 + nq + j;
      r.b[j] = (gn < N && gk < K) ? B[int64_t(gk) * pr.b_rs + gn] : 0.0;
    }
  } else {  // k contiguous (or general): KD/EB threads per column
    constexpr int TPC = KD / EB;
    const int n = tid / TPC, kq = (tid % TPC) * EB, gn = n0 + n;
#pragma unroll
    for (int j = 0; j < EB; ++j) {
      const int gk = k0 + kq + j;
      r.b[j] = (gn < N && gk < K) ? B[int64_t(gk) * pr.b_rs + int64_t(gn) * pr.b_cs] : 0.0;
    }
  }
}

template <int TM, int TN, int KD>
__device__ __forceinline__ void stash_t(const GemmPair& pr, int tid, const TileRegs<TM, TN, KD>& r,
                                        double (*As)[TM + 16], double (*Bs)[TN + 16]) {
  constexpr int EA = TM * KD / 256, EB = TN * KD / 256;
  if (pr.a_cs == 1) {
    constexpr int TPR = KD / EA;
    const int m = tid / TPR, kq = (tid % TPR) * EA;
#pragma unroll
    for (int j = 0; j < EA; ++j) As[kq + j][m] = r.a[j];
  } else {
    constexpr int TPK = TM / EA;
    const int k = tid / TPK, mq = (tid % TPK) * EA;
#pragma unroll
    for (int j = 0; j < EA; ++j) As[k][mq + j] = r.a[j];
  }
  if (pr.b_cs == 1) {
    constexpr int TPK = TN / EB;
    const int k = tid / TPK, nq = (tid % TPK) * EB;
#pragma unroll
    for (int j = 0; j < EB; ++j) Bs[k][nq + j] = r.b[j];
  } else {
    constexpr int TPC = KD / EB;
    const int n = tid / TPC, kq = (tid % TPC) * EB;
#pragma unroll
    for (int j = 0; j < EB; ++j) Bs[kq + j][n] = r.b[j];
  }
}

template <int TM, int TN, int KD>
__global__ __launch_bounds__(256) void k_gemm_f64t(double* __restrict__ base, const GemmTile* __restrict__ tiles,
                                                   const GemmOut* __restrict__ outs,
                                                   const GemmPair* __restrict__ pairs, double alpha, double beta) {
  constexpr int WM = TM / 2, WN = TN / 2, FI = WM / 16, FJ = WN / 16;
  __shared__ double As[2][KD][TM + 16];
  __shared__ double Bs[2][KD][TN + 16];
  const GemmTile t = tiles[blockIdx.x];
  const GemmOut o = outs[t.out];
  const int m0 = t.tm * TM, n0 = t.tn * TN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * WM, wn = (wave & 1) * WN;
  v4d acc[FI][FJ];
#pragma unroll
  for (int i = 0; i < FI; ++i)
#pragma unroll
    for (int j = 0; j < FJ; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};

  int p = o.pair_begin;
  while (p < o.pair_end && pairs[p].K <= 0) ++p;
  if (p < o.pair_end) {
    GemmPair pr = pairs[p];
    int k0 = 0, buf = 0;
    TileRegs<TM, TN, KD> regs;
    fetch_t<TM, TN, KD>(base, pr, o.M, o.N, m0, n0, 0, tid, regs);
    stash_t<TM, TN, KD>(pr, tid, regs, As[0], Bs[0]);
    __syncthreads();
    // the step after (p, k0)
    auto advance = [&](int& q, int& k, GemmPair& g) {
      k += KD;
      if (k >= g.K) {
        k = 0;
        ++q;
        while (q < o.pair_end && pairs[q].K <= 0) ++q;
        if (q < o.pair_end) g = pairs[q];
      }
    };
    int np = p, nk = k0;
    GemmPair npr = pr;
    advance(np, nk, npr);
    bool more = np < o.pair_end;
    if (more) fetch_t<TM, TN, KD>(base, npr, o.M, o.N, m0, n0, nk, tid, regs);
    while (true) {
#pragma unroll
      for (int kk = 0; kk < KD; kk += 4) {
        const int kr = kk + (lane >> 4), c = lane & 15;
        double a[FI], b[FJ];
#pragma unroll
        for (int i = 0; i < FI; ++i) a[i] = As[buf][kr][wm + i * 16 + c];
#pragma unroll
        for (int j = 0; j < FJ; ++j) b[j] = Bs[buf][kr][wn + j * 16 + c];
#pragma unroll
        for (int i = 0; i < FI; ++i)
#pragma unroll
          for (int j = 0; j < FJ; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
      }
      if (!more) break;
      // regs hold step (np, nk): into the other buffer (its last readers
      // passed the barrier of the previous step), then fetch the step after
      stash_t<TM, TN, KD>(npr, tid, regs, As[buf ^ 1], Bs[buf ^ 1]);
      GemmPair cur = npr;
      int q = np, k = nk;
      advance(q, k, npr);
      more = q < o.pair_end;
      np = q;
      nk = k;
      if (more) fetch_t<TM, TN, KD>(base, npr, o.M, o.N, m0, n0, nk, tid, regs);
      (void)cur;
      __syncthreads();
      buf ^= 1;
      // the step in buffer buf is complete; one more compute round if it
      // was the last one loaded
      if (!more) {
#pragma unroll
        for (int kk = 0; kk < KD; kk += 4) {
          const int kr = kk + (lane >> 4), c = lane & 15;
          double a[FI], b[FJ];
#pragma unroll
          for (int i = 0; i < FI; ++i) a[i] = As[buf][kr][wm + i * 16 + c];
#pragma unroll
          for (int j = 0; j < FJ; ++j) b[j] = Bs[buf][kr][wn + j * 16 + c];
#pragma unroll
          for (int i = 0; i < FI; ++i)
#pragma unroll
            for (int j = 0; j < FJ; ++j)
              acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        break;
      }
    }
  }
  double* C = base + o.c_off;
#pragma unroll
  for (int i = 0; i < FI; ++i)
#pragma unroll
    for (int j = 0; j < FJ; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm + i * 16 + 4 * r + (lane >> 4);
        const int col = n0 + wn + j * 16 + (lane & 15);
        if (row < o.M && col < o.N) {
          double* pc = C + int64_t(row) * o.ldc + col;
          double v = alpha * acc[i][j][r];
          if (beta != 0.0) v += beta * (*pc);
          *pc = v;
        }
      }
}

__global__ __launch_bounds__(256) void k_gemm_split(double* __restrict__ base, const GemmTile* __restrict__ tiles,
                                                    const GemmOut* __restrict__ outs,
                                                    const GemmPair* __restrict__ pairs, double alpha,
                                                    double beta) {
  __shared__ float Ah[BK][LDP], Al[BK][LDP];
  __shared__ float Bh[BK][LDP], Bl[BK][LDP];
  const GemmTile t = tiles[blockIdx.x];
  const GemmOut o = outs[t.out];
  const int m0 = t.tm * BM, n0 = t.tn * BN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  v4d acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};

  for (int p = o.pair_begin; p < o.pair_end; ++p) {
    const GemmPair pr = pairs[p];
    const double* A = base + pr.a_off;
    const double* B = base + pr.b_off;
    for (int k0 = 0; k0 < pr.K; k0 += BK) {
      load_tiles<float>(A, B, pr, o.M, o.N, m0, n0, k0, tid, Ah, Bh, Al, Bl);
      __syncthreads();
      v4f hh[2][2], cr[2][2];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          hh[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
          cr[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
        }
#pragma unroll
      for (int kk = 0; kk < BK; kk += 4) {
        const int kr = kk + (lane >> 4), c = lane & 15;
        float ah[2] = {Ah[kr][wm + c], Ah[kr][wm + 16 + c]};
        float al[2] = {Al[kr][wm + c], Al[kr][wm + 16 + c]};
        float bh[2] = {Bh[kr][wn + c], Bh[kr][wn + 16 + c]};
        float bl[2] = {Bl[kr][wn + c], Bl[kr][wn + 16 + c]};
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            hh[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ah[i], bh[j], hh[i][j], 0, 0, 0);
            cr[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ah[i], bl[j], cr[i][j], 0, 0, 0);
            cr[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(al[i], bh[j], cr[i][j], 0, 0, 0);
          }
      }
      // fold this K chunk into the FP64 accumulators
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) acc[i][j][r] += double(hh[i][j][r]) + double(cr[i][j][r]);
      __syncthreads();
    }
  }
  store_tile<false>(base + o.c_off, o, m0, n0, wm, wn, lane, acc, alpha, beta);
}

// Narrow outputs (N <= kRowsMaxN, summed K <= kRowsMaxK): an MFMA tile
// would be mostly padding there.  One thread per output row: B of every
// pair (at most kRowsMaxK x kRowsMaxN) is staged in LDS once per work-group,
// each thread streams its row of A (lanes read consecutive rows when A's row
// stride is 1) and keeps the N accumulators in registers.  Every element is
// one FP64 FMA chain over the pairs in order, k ascending.
__global__ __launch_bounds__(kRowsPerTask) void k_gemm_rows(double* __restrict__ base,
                                                             const RowTask* __restrict__ tasks,
                                                             const GemmOut* __restrict__ outs,
                                                             const GemmPair* __restrict__ pairs, double alpha,
                                                             double beta) {
  __shared__ double Bs[kRowsMaxK * kRowsMaxN];
  const RowTask t = tasks[blockIdx.x];
  const GemmOut o = outs[t.out];
  int kb = 0;
  for (int p = o.pair_begin; p < o.pair_end; ++p) {
    const GemmPair pr = pairs[p];
    const double* B = base + pr.b_off;
    for (int e = int(threadIdx.x); e < pr.K * o.N; e += blockDim.x) {
      const int k = e / o.N, n = e - k * o.N;
      Bs[(kb + k) * kRowsMaxN + n] = B[int64_t(k) * pr.b_rs + int64_t(n) * pr.b_cs];
    }
    kb += pr.K;
  }
  __syncthreads();
  const int m = t.row0 + int(threadIdx.x);
  if (m >= o.M) return;
  double acc[kRowsMaxN];
#pragma unroll
  for (int n = 0; n < kRowsMaxN; ++n) acc[n] = 0.0;
  kb = 0;
  for (int p = o.pair_begin; p < o.pair_end; ++p) {
    const GemmPair pr = pairs[p];
    const double* A = base + pr.a_off + int64_t(m) * pr.a_rs;
#pragma unroll 4
    for (int k = 0; k < pr.K; ++k) {
      const double a = A[int64_t(k) * pr.a_cs];
      const double* bk = Bs + (kb + k) * kRowsMaxN;
#pragma unroll
      for (int n = 0; n < kRowsMaxN; ++n)
        if (n < o.N) acc[n] = fma(a, bk[n], acc[n]);
    }
    kb += pr.K;
  }
  double* C = base + o.c_off + int64_t(m) * o.ldc;
#pragma unroll
  for (int n = 0; n < kRowsMaxN; ++n)
    if (n < o.N) {
      double v = alpha * acc[n];
      if (beta != 0.0) v += beta * C[n];
      C[n] = v;
    }
}

void gemm_rows(hipStream_t s, double* base, const RowTask* tasks, int ntasks, const GemmOut* outs,
               const GemmPair* pairs, double alpha, double beta) {
  if (ntasks <= 0) return;
  hipLaunchKernelGGL(k_gemm_rows, dim3(ntasks), dim3(kRowsPerTask), 0, s, base, tasks, outs, pairs, alpha, beta);
}

void gemm_f64(hipStream_t s, double* base, const GemmTile* tiles, int ntiles, const GemmOut* outs,
              const GemmPair* pairs, double alpha, double beta) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_gemm_f64, dim3(ntiles), dim3(256), 0, s, base, tiles, outs, pairs, alpha, beta);
}

void gemm_f64_variant(hipStream_t s, int variant, double* base, const GemmTile* tiles, int ntiles,
                      const GemmOut* outs, const GemmPair* pairs, double alpha, double beta) {
  if (ntiles <= 0) return;
  switch (variant) {
    case kTile64:
      hipLaunchKernelGGL(k_gemm_f64, dim3(ntiles), dim3(256), 0, s, base, tiles, outs, pairs, alpha, beta);
      break;
    case kTile64db:
      hipLaunchKernelGGL((k_gemm_f64t<64, 64, 16>), dim3(ntiles), dim3(256), 0, s, base, tiles, outs, pairs, alpha, beta);
      break;
    case kTile128x64:
      hipLaunchKernelGGL((k_gemm_f64t<128, 64, 16>), dim3(ntiles), dim3(256), 0, s, base, tiles, outs, pairs, alpha,
                         beta);
      break;
    case kTile64x128:
      hipLaunchKernelGGL((k_gemm_f64t<64, 128, 16>), dim3(ntiles), dim3(256), 0, s, base, tiles, outs, pairs, alpha,
                         beta);
      break;
    case kTile128:
      hipLaunchKernelGGL((k_gemm_f64t<128, 128, 16>), dim3(ntiles), dim3(256), 0, s, base, tiles, outs, pairs, alpha,
                         beta);
      break;
    case kTile64k32:
      hipLaunchKernelGGL((k_gemm_f64t<64, 64, 32>), dim3(ntiles), dim3(256), 0, s, base, tiles, outs, pairs, alpha,
                         beta);
      break;
    default:
      break;
  }
}

void gemm_split(hipStream_t s, double* base, const GemmTile* tiles, int ntiles, const GemmOut* outs,
                const GemmPair* pairs, double alpha, double beta) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_gemm_split, dim3(ntiles), dim3(256), 0, s, base, tiles, outs, pairs, alpha, beta);
}

// Scratch of the allocation-free contraction (packed operands, split-K
// partials, launch descriptors).  Every GPU the library runs on holds its
// own copy for the life of the process.  In host code this name stands for
// the variable, not for a device address: host code uses only the address
// hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

hipError_t noalloc_scratch_address(int device, void** out) {
  static std::mutex mu;
  static std::map<int, void*> addr;
  std::lock_guard<std::mutex> lk(mu);
  auto it = addr.find(device);
  if (it != addr.end()) {
    *out = it->second;
    return hipSuccess;
  }
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  addr[device] = p;
  *out = p;
  return hipSuccess;
}

}  // namespace kern
}  // namespace qtt
