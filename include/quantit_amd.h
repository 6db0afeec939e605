/*
 * C ABI of the MI355X block-sparse tensor library (libquantit_amd.so).
 *
 * The entry points are the operations a foreign-function binding of QuantiT's
 * block-sparse DMRG path calls: btensor construction and inspection, the
 * block contraction, the block SVD/eigh with truncation and the DMRG driver.
 * Plain pointers and sizes only.  Every function returns a status
 * (QTT_OK = 0); on failure qtt_last_error() returns the message of the
 * exception the reference would have thrown, and the status names its type.
 * Block data lives in the device arena of a context; tensors of one context
 * are used from one host thread at a time.
 */
#ifndef QUANTIT_AMD_H
#define QUANTIT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QTT_OK 0
#define QTT_INVALID_ARGUMENT 1 /* std::invalid_argument */
#define QTT_OUT_OF_RANGE 2     /* std::out_of_range */
#define QTT_DOMAIN_ERROR 3     /* std::domain_error */
#define QTT_LOGIC_ERROR 4      /* std::logic_error */
#define QTT_TORCH_CHECK 5      /* c10::Error from TORCH_CHECK */
#define QTT_OUT_OF_MEMORY 6    /* device arena exhausted */
#define QTT_HIP_ERROR 7        /* HIP runtime failure */
#define QTT_RUNTIME_ERROR 8

#define QTT_PRECISION_FP64 0       /* FP64 MFMA contractions */
#define QTT_PRECISION_SPLIT_FP32 1 /* hi/lo split FP32 MFMA, FP64 recombination */

typedef struct qtt_context qtt_context;
typedef struct qtt_btensor qtt_btensor;

/* Message of the last failed call on this thread. */
const char* qtt_last_error(void);
int qtt_version(void);

/* ---- context: device, HIP stream and HBM arena ------------------------- */
int qtt_context_create(int device, size_t arena_bytes, qtt_context** out);
int qtt_context_destroy(qtt_context* ctx);
int qtt_context_set_precision(qtt_context* ctx, int precision);
/* 1: one GEMM launch per block pair (reference-shaped baseline) */
int qtt_context_set_per_block(qtt_context* ctx, int enable);
/* 1 (default): independent work of one call (the narrow-sector solves of a
   block SVD beside the wide ones) may run on a second stream; 0: one stream.
   No counterpart in the reference.  Processes that drive several contexts at
   once (concurrent solves) set 0: their streams already share the device. */
int qtt_context_set_overlap(qtt_context* ctx, int enable);
int qtt_context_set_profiling(qtt_context* ctx, int enable);
int qtt_context_sync(qtt_context* ctx);
int qtt_context_memory(qtt_context* ctx, size_t* used, size_t* peak, size_t* capacity);
/* Writes "name ms launches\n" lines of the accumulated per-kernel times. */
int qtt_context_kernel_times(qtt_context* ctx, char* buf, size_t buflen, int reset);
/* Block SVD/eigh health since the context was created: calls, the largest
   number of sweeps (Gram iterations + Jacobi sweeps) one call needed, and
   sectors whose Jacobi sweeps reached the sweep cap while still rotating
   (their singular values are then less accurate than the tolerance).  No
   counterpart in the reference, whose LAPACK svd raises on non-convergence. */
int qtt_context_svd_stats(qtt_context* ctx, int64_t* calls, int64_t* max_sweeps, int64_t* unconverged);
/* Block SVD model FLOPs (sum over sectors of 14 m n^2 + 8 n^3, m >= n) since
   the context was created: all sectors, and the sectors whose rotations ran
   as multi-work-group block-Jacobi rounds (the wide ones). */
int qtt_context_svd_flops(qtt_context* ctx, double* all, double* wide);
/* Eigensolver of the DMRG two-site update.  max_krylov = 0 (default): the
   reference's one-step Lanczos with its closed-form 2x2 solve
   (details::one_step_lanczos + eig2x2Mat, sources/dmrg.cpp:543-608).
   max_krylov in [2, 64]: k-step Lanczos with full FP64 reorthogonalisation
   and an FP64 tridiagonal eigensolve on the device, stopping when the
   residual |H theta - E theta| <= tol |E|, at an invariant subspace or at
   max_krylov vectors (2 reproduces the one-step update).  No counterpart in
   the reference; it lets a solve converge to the energy tolerance instead
   of improving by one Lanczos step per update. */
int qtt_context_set_lanczos(qtt_context* ctx, int max_krylov, double tol);
/* k-step mode: the matvec of an update is captured once as a HIP graph and
   replayed for its later Krylov vectors (1, the default) or run eagerly (0).
   The results are the same bits either way.  No counterpart in the
   reference. */
int qtt_context_set_graphs(qtt_context* ctx, int enable);
/* k-step mode: two-site updates run, Krylov vectors built and matvecs run as
   graph replays so far. */
int qtt_context_lanczos_stats(qtt_context* ctx, int64_t* updates, int64_t* vectors, int64_t* graph_replays);

/* ---- btensor ------------------------------------------------------------
 * Replaces btensor(index_list sections_by_dim, any_quantity_vector c_vals,
 * index_list section_sizes, any_quantity sel_rule)
 * (include/blockTensor/btensor.h:128-133) plus the filling of its blocks.
 * nsec[rank]: sections per dimension; sec_sizes[sum nsec]; cvals[sum nsec *
 * ngroups] and sel[ngroups]: charges as int16 tuples; group_mod[ngroups]: 0
 * for Z, N for C<N>.  block_idx[nblocks * rank] lists the blocks in
 * lexicographic order; data holds them concatenated, each row-major. */
int qtt_btensor_create(qtt_context* ctx, int rank, const int64_t* nsec, const int64_t* sec_sizes, int ngroups,
                       const uint16_t* group_mod, const int16_t* cvals, const int16_t* sel, int64_t nblocks,
                       const int32_t* block_idx, const double* data, qtt_btensor** out);
int qtt_btensor_destroy(qtt_btensor* t);
/* rank, number of groups, total section count, block count, element count */
int qtt_btensor_info(const qtt_btensor* t, int* rank, int* ngroups, int64_t* total_sections, int64_t* nblocks,
                     int64_t* numel);
/* Structure readback: nsec[rank], sec_sizes[total_sections],
 * cvals[total_sections * ngroups], sel[ngroups], group_mod[ngroups]. */
int qtt_btensor_structure(const qtt_btensor* t, int64_t* nsec, int64_t* sec_sizes, int16_t* cvals, int16_t* sel,
                          uint16_t* group_mod);
/* Block readback (synchronising): block_idx[nblocks * rank], data[numel]. */
int qtt_btensor_blocks(const qtt_btensor* t, int32_t* block_idx, double* data);

/* ---- structural operations -------------------------------------------------- */
/* btensor::permute (btensor.h:555, sources/btensor.cpp:1754-1802) */
int qtt_permute(const qtt_btensor* t, const int64_t* perm, qtt_btensor** out);
/* btensor::reshape (btensor.h:569, sources/btensor.cpp:2986-3024) */
int qtt_reshape(const qtt_btensor* t, int ngroups, const int64_t* index_groups, qtt_btensor** out);
/* btensor::conj (btensor.h:649, sources/btensor.cpp:2156) */
int qtt_conj(const qtt_btensor* t, qtt_btensor** out);

/* ---- contraction ------------------------------------------------------------
 * btensor::tensordot (include/blockTensor/btensor.h:623) and the free
 * tensordot(l, r, dl, dr) (:1230-1234), bound in Python at
 * python_binding/quantt.cpp:581-597. */
int qtt_tensordot(const qtt_btensor* a, const qtt_btensor* b, int ndims, const int64_t* dims_a,
                  const int64_t* dims_b, qtt_btensor** out);
/* Host-counted algorithmic FLOPs (sum of 2 M N K) of the last tensordot on this thread. */
double qtt_last_tensordot_flops(void);
/* No counterpart in the reference: qtt_tensordot for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The result is written into `out`, an
 * existing tensor of the same context with exactly the result's structure
 * (sections, charges, selection rule, block list in order), for instance a
 * previous result of the same contraction; it must not share storage with a
 * or b.  Packed operands, split-K partials and launch descriptors live in a
 * static device scratch of 128 MiB per GPU; a contraction that needs more
 * fails with QTT_INVALID_ARGUMENT before any launch.  The call allocates
 * nothing, records no profiling events and always uses the grouped kernels.
 * Every call on one GPU shares that scratch: calls that overlap on one GPU
 * (several contexts driven from several threads) race. */
int qtt_tensordot_into(const qtt_btensor* a, const qtt_btensor* b, int ndims, const int64_t* dims_a,
                       const int64_t* dims_b, qtt_btensor* out);

/* The FP64 contraction GEMM has tile variants (64x64 single-buffered, and
   64x64, 128x64, 64x128, 128x128 with double-buffered LDS); they give the
   same bits.  A contraction's shape class (work-weighted mean block size and
   tile count) picks its variant from a process-wide table that is filled by
   timing every variant the first time a class is met; the table persists as
   JSON through these calls.  No counterpart in the reference. */
int qtt_gemm_tuning_save(const char* path);
int qtt_gemm_tuning_load(const char* path);
int qtt_gemm_tuning_size(int64_t* entries);

/* ---- elementwise ------------------------------------------------------------ */
/* broadcast mul (btensor.h:547, sources/btensor.cpp:1750) */
int qtt_mul(const qtt_btensor* x, const qtt_btensor* y, qtt_btensor** out);
/* add(other, alpha) (btensor.h:462, sources/btensor.cpp:2666) */
int qtt_add(const qtt_btensor* x, const qtt_btensor* y, double alpha, qtt_btensor** out);
/* rank-0 tensordot(x, y.conj(), all dims) */
int qtt_full_dot(const qtt_btensor* x, const qtt_btensor* y, double* out);

/* ---- block linear algebra (include/blockTensor/LinearAlgebra.h) --------- */
/* svd(const btensor&, size_t split) (LinearAlgebra.h:87) */
int qtt_svd(const qtt_btensor* t, size_t split, qtt_btensor** U, qtt_btensor** d, qtt_btensor** V);
/* svd(const btensor&, size_t split, tol, min_size, max_size, pow) (LinearAlgebra.h:115-116) */
int qtt_svd_truncated(const qtt_btensor* t, size_t split, double tol, size_t min_size, size_t max_size, double pow,
                      qtt_btensor** U, qtt_btensor** d, qtt_btensor** V);
/* eigh(const btensor&, size_t split) (LinearAlgebra.h:167) */
int qtt_eigh(const qtt_btensor* t, size_t split, qtt_btensor** d, qtt_btensor** U);
/* truncate(std::tuple<U, d, V>&&, max, min, tol, pow) (LinearAlgebra.h:244-247) */
int qtt_truncate(const qtt_btensor* U, const qtt_btensor* d, const qtt_btensor* V, size_t max_size, size_t min_size,
                 double tol, double pow, qtt_btensor** U_out, qtt_btensor** d_out, qtt_btensor** V_out);
/* svd(const btensor&, BOOL some, BOOL compute_uv) of a rank-2 tensor
 * (LinearAlgebra.h:74, sources/btensor_linalg.cpp:390-502).  some = 0 asks for
 * full square unitaries; the reference then sizes d, U and V by max(m, n) per
 * sector while the sector factors have min(m, n) values, and compute_uv = 0
 * (torch::svd's m x m and n x n zero matrices) is no better, so in those
 * modes only sectors with m = n are accepted (there every form has the
 * economy structure) and a rectangular sector gives QTT_INVALID_ARGUMENT.
 * compute_uv = 0 returns U and V filled with zeros. */
int qtt_svd_matrix(const qtt_btensor* t, int some, int compute_uv, qtt_btensor** U, qtt_btensor** d,
                   qtt_btensor** V);
/* eigh(const btensor&, size_t split, tol, min_size, max_size, pow)
 * (LinearAlgebra.h:180-181, sources/btensor_linalg.cpp:816-823): eigh then
 * the eigen truncate below.  The reference's default pow here is 1. */
int qtt_eigh_truncated(const qtt_btensor* t, size_t split, double tol, size_t min_size, size_t max_size, double pow,
                       qtt_btensor** d, qtt_btensor** U);
/* truncate(btensor&& e, btensor&& S, max, min, tol, pow) (LinearAlgebra.h:242-243,
 * sources/btensor_linalg.cpp:786-805): truncate_impl with the eigenvalues as
 * weights and the eigenvectors S as the one unitary.  Eigenvalues are
 * ascending within a sector, so a sector is kept whole or erased whole. */
int qtt_truncate_eigh(const qtt_btensor* e, const qtt_btensor* S, size_t max_size, size_t min_size, double tol,
                      double pow, qtt_btensor** e_out, qtt_btensor** S_out);

/* ---- DMRG ----------------------------------------------------------------
 * dmrg(bMPO&, bMPS&, const dmrg_options&, dmrg_logger&) (include/dmrg.h:39-40,
 * bound at python_binding/algorithm.cpp:46-62).  options mirrors
 * dmrg_options (include/dmrg_options.h:15-57).  mps[] is replaced in place by
 * the optimised state (the old handles are destroyed).  Per-sweep energies,
 * wall times (ms) and largest bond are written for up to max_log sweeps. */
typedef struct qtt_dmrg_options {
  double cutoff;
  double convergence_criterion;
  uint64_t maximum_bond;
  uint64_t minimum_bond;
  uint64_t maximum_iterations;
} qtt_dmrg_options;
int qtt_dmrg(int L, qtt_btensor* const* mpo, qtt_btensor** mps, const qtt_dmrg_options* options, double* energy,
             int max_log, double* sweep_energies, double* sweep_ms, int64_t* max_bond, int* n_sweeps,
             double* contract_flops, double* svd_flops);
/* The same with the state's orthogonality centre (bMPS::orthogonality_center)
 * at oc instead of 0: sweeps start toward the far end, a centre on the last
 * site starts one site to the left, and the centre is moved back to oc at
 * the end (details::dmrg_impl, sources/dmrg.cpp:224-236, :257-266). */
/* dmrg(bMPO&, bMPS&, options, dmrg_logger&) with the logger's hooks
 * (include/dmrg_logger.h:27-108; called as sources/dmrg.cpp:239-268 does):
 * init(user, options) before the first sweep; it_log_all(user, iteration,
 * energy, L, state) after every sweep; end_log_all(user, iteration, energy,
 * L, state) once the orthogonality centre is back at oc.  The state handles
 * are lent for the duration of the callback (the callback may pass them to
 * any entry point that does not destroy them, and must not keep them).
 * Null function pointers are skipped.  Otherwise as qtt_dmrg_from_center. */
typedef struct qtt_dmrg_logger {
  void* user;
  void (*init)(void* user, const qtt_dmrg_options* options);
  void (*it_log_all)(void* user, size_t iteration, double energy, int L, qtt_btensor* const* state);
  void (*end_log_all)(void* user, size_t iteration, double energy, int L, qtt_btensor* const* state);
} qtt_dmrg_logger;
int qtt_dmrg_logged(int L, qtt_btensor* const* mpo, qtt_btensor** mps, int oc, const qtt_dmrg_options* options,
                    const qtt_dmrg_logger* logger, double* energy);
int qtt_dmrg_from_center(int L, qtt_btensor* const* mpo, qtt_btensor** mps, int oc, const qtt_dmrg_options* options,
                         double* energy, int max_log, double* sweep_energies, double* sweep_ms, int64_t* max_bond,
                         int* n_sweeps, double* contract_flops, double* svd_flops);
/* details::hamil2site_times_state (sources/dmrg.cpp:520-531): the two-site
 * matvec L.theta.W12.R (three tensordots). */
int qtt_hamil2site_times_state(const qtt_btensor* theta, const qtt_btensor* W12, const qtt_btensor* L,
                               const qtt_btensor* R, qtt_btensor** out);
/* details::one_step_lanczos (sources/dmrg.cpp:584-608): Hpsi receives the
 * normalised Lanczos vector (H theta - a0 theta)/b (undivided when |b| <
 * 1e-15); a0 = <theta|H|theta>, a1 = <Hpsi|H|Hpsi>, b the residual norm. */
int qtt_one_step_lanczos(const qtt_btensor* theta, const qtt_btensor* W12, const qtt_btensor* L,
                         const qtt_btensor* R, qtt_btensor** Hpsi, double* a0, double* a1, double* b);
/* details::eig2x2Mat (sources/dmrg.cpp:543-572) evaluated on the context's
 * device: lowest eigenvalue E of [[a0, b], [b, a1]] and its eigenvector
 * (o, n).  A NaN gives QTT_LOGIC_ERROR, as the reference's std::logic_error. */
int qtt_eig2x2(qtt_context* ctx, double a0, double a1, double b, double* E, double* o, double* n);

/* bMPS::move_oc (sources/MPT.cpp:74-111): moves the orthogonality centre of
 * the L tensors mps[] from oc to target with untruncated block SVDs; the
 * handles are replaced in place. */
int qtt_mps_move_oc(int L, qtt_btensor** mps, int oc, int target);
/* Preparation of a random bMPS before dmrg (tests/DMRG_heisenberg_test.h:34-37):
 * mps[0] /= sqrt(<psi|psi>), move_oc(L-1), move_oc(0).  Handles replaced. */
int qtt_mps_prepare(int L, qtt_btensor** mps);
/* contract(bMPS, bMPS) and contract(bMPS, bMPS, bMPO) (sources/MPT.cpp:211-292) */
int qtt_contract_norm(int L, qtt_btensor* const* a, qtt_btensor* const* b, double* out);
int qtt_contract_energy(int L, qtt_btensor* const* a, qtt_btensor* const* b, qtt_btensor* const* mpo, double* out);

#ifdef __cplusplus
}
#endif

#endif /* QUANTIT_AMD_H */
