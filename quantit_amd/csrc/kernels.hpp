// Host-side launchers of the gfx950 kernels.  Every device pointer is the
// context arena base plus an element (double) offset, so descriptor lists are
// plain integer records.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace qtt {
namespace kern {

// ---- grouped block GEMM ---------------------------------------------------
// C_out = alpha * sum_p A_p * B_p + beta * C_out for every output block, the
// pairs of one output block taken in list order (the reference's mm +
// addmm_ order, sources/btensor.cpp:2095-2102).  A_p[m,k] =
// base[a_off + m*a_rs + k*a_cs], B_p[k,n] = base[b_off + k*b_rs + n*b_cs];
// C is row-major M x N at c_off with row pitch ldc.
struct GemmOut {
  int64_t c_off;
  int32_t M, N;
  int32_t pair_begin, pair_end;
  int32_t ldc, pad;  // row pitch of C (N for a packed block)
};
struct GemmPair {
  int64_t a_off, b_off;
  int64_t a_rs, a_cs, b_rs, b_cs;
  int32_t K, pad;
};
struct GemmTile {
  int32_t out;
  int16_t tm, tn;  // tile coordinates inside the output block
};
constexpr int kGemmTile = 64;

void gemm_f64(hipStream_t s, double* base, const GemmTile* tiles, int ntiles, const GemmOut* outs,
              const GemmPair* pairs, double alpha, double beta);
// Tile variants of the FP64 grouped GEMM (same bits as gemm_f64; a tile's
// tm, tn count in units of the variant's TM x TN).  kTile64 is gemm_f64.
// kTile64k32 is the double-buffered 64 x 64 tile with 32-deep k steps.
enum GemmVariant : int {
  kTile64 = 0,
  kTile64db = 1,
  kTile128x64 = 2,
  kTile64x128 = 3,
  kTile128 = 4,
  kTile64k32 = 5,
  kGemmVariants = 6
};
constexpr int kVariantTM[kGemmVariants] = {64, 64, 128, 64, 128, 64};
constexpr int kVariantTN[kGemmVariants] = {64, 64, 64, 128, 128, 64};
void gemm_f64_variant(hipStream_t s, int variant, double* base, const GemmTile* tiles, int ntiles,
                      const GemmOut* outs, const GemmPair* pairs, double alpha, double beta);

// Size of the file-scope device scratch (kernels/gemm.hip) that the
// allocation-free contraction takes its packed operands, split-K partials
// and descriptors from.
constexpr size_t kNoAllocScratchBytes = size_t(128) << 20;
// Device address of that scratch on `device` (which must be current):
// hipGetSymbolAddress on the first call per device, cached afterwards.
hipError_t noalloc_scratch_address(int device, void** out);
// hi/lo split FP32 MFMA: each operand x = hi + lo (hi = fp32(x), lo =
// fp32(x - hi)); products hi*hi + hi*lo + lo*hi accumulate in FP32 per K
// chunk and are folded into FP64 accumulators after every chunk.
void gemm_split(hipStream_t s, double* base, const GemmTile* tiles, int ntiles, const GemmOut* outs,
                const GemmPair* pairs, double alpha, double beta);
// Output blocks with at most kRowsMaxN columns and a short summed K: one
// thread per output row (B staged in LDS), the pairs in order, k ascending
// (one FP64 FMA chain per element).  A task is kRowsPerTask consecutive rows
// of one output block.
constexpr int kRowsMaxN = 16, kRowsMaxK = 128, kRowsPerTask = 256;
struct RowTask {
  int32_t out, row0;
};
inline int rows_per_task(int) { return kRowsPerTask; }
void gemm_rows(hipStream_t s, double* base, const RowTask* tasks, int ntasks, const GemmOut* outs,
               const GemmPair* pairs, double alpha, double beta);

// ---- data movement ----------------------------------------------------------
// Row-major copy of a permuted view: dst element with multi-index j over
// dims[0..rank) (row-major) reads src[sum_k j_k * sstride[k]].
struct PermDesc {
  int64_t src_off, dst_off, numel;
  int32_t rank, pad;
  int32_t dims[8];
  int64_t sstride[8];
};
struct ChunkTask {
  int32_t desc;
  int32_t chunk;
};
// Elements per work-group task of the element-wise kernels: 8 per thread of
// a 256-thread group, so a task is a few dependent memory round trips and a
// launch spreads over many work-groups.
constexpr int kChunk = 2048;
void permute_copy(hipStream_t s, double* base, const PermDesc* descs, const ChunkTask* tasks, int ntasks);

// 2-D strided copy dst[r*dst_rs + c*dst_cs] = alpha * src[r*src_rs + c*src_cs];
// src_off < 0 fills the destination with the value alpha.
struct Copy2D {
  int64_t src_off, dst_off;
  int64_t src_rs, src_cs, dst_rs, dst_cs;
  int32_t rows, cols;
  double alpha;
};
void copy2d(hipStream_t s, double* base, const Copy2D* descs, const ChunkTask* tasks, int ntasks);
struct SumParts {
  int64_t dst_off, src_off, n;
  int32_t parts, pad;
};
// dst = sum of `parts` consecutive arrays of n doubles at src (fixed order).
void sum_parts(hipStream_t s, double* base, const SumParts* descs, const ChunkTask* tasks, int ntasks);

// ---- elementwise / reductions ---------------------------------------------
// out = x (op) y with y broadcast: mode 0 y scalar, 1 y vector along the last
// dim (x viewed as rows x ncol), 2 y same shape.  op 0 = mul, 1 = div.
// For op 1 with mode 0 and min_abs > 0 the division is skipped (plain copy)
// when |y| < min_abs.
struct EwDesc {
  int64_t x_off, y_off, out_off, numel;
  int32_t ncol, mode;
};
void ew_binary(hipStream_t s, double* base, const EwDesc* descs, const ChunkTask* tasks, int ntasks, int op,
               double min_abs);
// out = a*x + b*y; x_off or y_off < 0 means that operand is absent.  a and b
// are read from device scalars when a_off/b_off >= 0, else the host values.
struct AxpbyDesc {
  int64_t x_off, y_off, out_off, numel;
};
void axpby(hipStream_t s, double* base, const AxpbyDesc* descs, const ChunkTask* tasks, int ntasks, int64_t a_off,
           double a, int64_t b_off, double b);
// Deterministic dot product over matched block pairs: one partial per chunk
// task, then a fixed-order tree reduction into base[out_off].  With sqrt_out
// the square root is stored instead.
struct DotDesc {
  int64_t x_off, y_off, numel;
};
void dot(hipStream_t s, double* base, const DotDesc* descs, const ChunkTask* tasks, int ntasks,
         int64_t partial_off, int64_t out_off, bool sqrt_out);
// eig2x2Mat (sources/dmrg.cpp:543-572) on device scalars a0, a1, b; writes
// E, o, n to out[0..3) and sets flags[0] = 1 on NaN.
void eig2x2(hipStream_t s, double* base, int64_t a0_off, int64_t a1_off, int64_t b_off, int64_t out_off,
            int* flags);
// d /= sqrt(sum d^2) over n contiguous values.
void normalize(hipStream_t s, double* base, int64_t off, int64_t n);

// ---- k-step Lanczos (kernels/lanczos.hip) -----------------------------------
// Krylov vectors are the rows of V (row pitch ld doubles, n used).
constexpr int kMaxKrylov = 64;
constexpr int kKrylovPerThread = 8;
constexpr int kKrylovChunk = 256 * kKrylovPerThread;
// out[i] = <V[i], w> for i < m (partials: m * ceil(n / kKrylovChunk) doubles).
void krylov_dot(hipStream_t s, double* base, int64_t v_off, int64_t ld, int m, int64_t w_off, int64_t n,
                int64_t partial_off, int64_t out_off);
// dst = src + sign * sum_i coef[i] V[i] (src_off < 0: no src term), divided by
// sqrt(base[scale_off]) when scale_off >= 0.
void krylov_axpy(hipStream_t s, double* base, int64_t v_off, int64_t ld, int m, int64_t coef_off, double sign,
                 int64_t src_off, int64_t dst_off, int64_t n, int64_t scale_off);
// alpha[j] = c1[j] + c2[j], beta[j] = sqrt(nrm2) into T (alpha T[0..K), beta
// T[K..2K)), then the lowest eigenpair of the leading (j+1)-square
// tridiagonal: out = {eigenvalue, residual beta[j] |y[j]|, beta[j], y...}.
void lanczos_tridiag(hipStream_t s, double* base, int64_t t_off, int K, int j, int64_t c1_off, int64_t c2_off,
                     int64_t nrm2_off, int64_t out_off, int* flags);

// ---- block SVD ----------------------------------------------------------
// Sector work matrices: Y (ell x k, column-major, ld = ell) is orthogonalised
// in place by block one-sided Jacobi; W (k x k, column-major) accumulates the
// rotations.  See svd_jacobi.hip.
struct SvdSector {
  int64_t y_off, w_off;     // work matrices (w_off < 0: no right rotation kept)
  int64_t u_off, v_off;     // outputs: U (m x k) and V (n x k) row-major
  int64_t s_off;            // singular values (k), descending
  int32_t m, n, ell, k;     // ell = max(m,n), k = min(m,n)
  int32_t nb;               // number of column blocks (even)
  int32_t transposed;       // 1 when Y = A^T (m < n)
  int32_t flag_idx;         // rotation counter slot
  int32_t eig_shift;        // 1: eigh via shifted SVD (A + c I), values written ascending
  int64_t shift_off;        // device double holding c (eigh mode)
};
constexpr int kSvdBlock = 16;  // columns per Jacobi block
void svd_init_w(hipStream_t s, double* base, const SvdSector* secs, int nsec);
// One round of the block Jacobi tournament; max_ell = largest row count
// among the sectors (picks the LDS-resident kernel when the block pair fits).
// cross_after0: rounds after the first of a sweep rotate only the pairs
// across the two blocks (the first round takes every pair of its blocks).
// skip (may be null): the previous sweep's counters; a sector whose entry is
// 0 has converged and its work-groups return at once.  A work item carries
// the sector fields a round needs, so a work-group starts with one load.
struct RoundWork {
  int64_t y_off, w_off;
  int32_t ell, nb, flag_idx, pair;
};
void svd_round(hipStream_t s, double* base, const RoundWork* work, int nwork, int round, int* counters,
               const int* skip, double tol, int inner_sweeps, int max_ell, int cross_after0);
// norms + ranks (one work-group per sector) and the tiled transposed scatter
// of the (sector, part, row0, col0) tasks into U / V.
void svd_finalize(hipStream_t s, double* base, const SvdSector* secs, const SvdSector* host_secs, int nsec,
                  const int4* tasks, int ntasks);
// Orthonormal completion of the Z-side columns (Z: ell x k row-major, values
// S descending) whose singular value is numerically zero, for columns
// t < limit only: a truncation that keeps `limit` columns never reads the
// others.  Scratch: v (ell doubles) and p (k doubles).
struct CompleteTask {
  int64_t z_off, s_off, v_off, p_off;
  int32_t ell, k, limit, pad;
};
void svd_complete(hipStream_t s, double* base, const CompleteTask* tasks, int n);
// eigh mode: c = ||A||_F per sector, stored at shift_off, and Y += c I.
void svd_shift(hipStream_t s, double* base, const SvdSector* secs, int nsec);
void fill_i32(hipStream_t s, int32_t* p, int n, int32_t v);

// Single-work-group one-sided Jacobi with the whole sector in LDS
// (descriptors: see SmallJacobi in svd_jacobi.hip).
struct SmallJacobi {
  int64_t y_off, w_off;  // Y (ell x k, column-major, ld ldy), W (ldw x ldw, column-major)
  int32_t ell, k, ldy, ldw;
  int32_t flag_idx;       // flags[flag_idx]: result count of the kernel (-1: none)
  int32_t aux_idx;        // flags[aux_idx]: second count (gram_eig; -1: none)
};
void jacobi_small(hipStream_t s, double* base, const void* descs, int n, int* flags, double tol, int max_sweeps,
                  size_t lds_bytes);
size_t jacobi_small_lds_limit();
// Eigenvectors of k x k Gram matrices (SmallJacobi with ell = k): W written
// k x k at w_off with leading dimension ldw; flags[flag_idx] = pairs of G
// failing the orthogonality test (0: W = I).  precond: pivoted-Cholesky
// preconditioning.
constexpr int kGramEigMaxK = 140;
void gram_eig(hipStream_t s, double* base, const void* descs, int n, int* flags, double tol, int max_sweeps,
              int precond, size_t lds_bytes);
size_t gram_eig_lds_bytes(int k);
// First-order Gram-Schmidt factor T (k x k at w_off) from the Gram matrix at
// y_off (SmallJacobi with ell = k).
void gs_step(hipStream_t s, double* base, const void* descs, int n, int kmax);

// ---- Gram eigenvectors for wide sectors (G in device memory) ---------------
// Orthogonality test of G (SmallJacobi: y_off = G, k, ldy; flag_idx/aux_idx
// as for gram_eig).
void gram_check(hipStream_t s, const double* base, const void* descs, int n, int* flags, double tol, int kmax);
struct CholPanel {
  int64_t g_off;      // G (k x k, ld k), updated between panels
  int64_t l_off;      // L (k x k, ld k, zeroed), rows in the original order
  int64_t lp_off;     // panel copy (k x nb)
  int64_t state_off;  // ints: [0] rank (k while running), [1 + i] index i pivoted
  int32_t k, pad;
};
// One panel of nb columns starting at j0; LDS kmax * nb doubles.
void pchol_panel(hipStream_t s, double* base, const void* descs, int n, int j0, int nb, int kmax);
// state of every factorisation: [0] = k (running), [1 .. k] = 0 (none pivoted)
void pchol_init(hipStream_t s, double* base, const void* descs, int n);
constexpr int kPcholExtra = 2;  // panel LDS per index besides the columns: running diagonal, flag
size_t pchol_lds_limit();
struct GramVec {
  int64_t x_off, w_off, scratch_off;  // scratch: 3k doubles
  int32_t k, ldx, ldw, pad;
};
void gram_vectors(hipStream_t s, double* base, const void* descs, int n, int kmax);  // largest dynamic LDS per work-group (bytes)

// Global truncation (sources/btensor_linalg.cpp:657-755,
// sources/LinearAlgebra.cpp:57-75): singular values of nsec sectors stored
// descending at s_off[i] with k[i] entries, s_pos[i] = k[0] + ... + k[i-1];
// writes kept[i].
void truncate_select(hipStream_t s, double* base, const int64_t* s_offs, const int32_t* ks, const int64_t* s_pos,
                     int nsec, int64_t scratch_off, int64_t total, double tol, double pow, int64_t min_size,
                     int64_t max_size, int32_t* kept, double* thr_out);

}  // namespace kern
}  // namespace qtt
