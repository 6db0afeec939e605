// Block-sparse tensor with host-side sector bookkeeping and device-resident
// block data.  Mirrors the state of QuantiT's btensor
// (include/blockTensor/btensor.h:820-837): sections per dimension with sizes
// and conserved charges, a selection rule, and a lexicographically ordered
// block list.  All blocks of one tensor live contiguously (row-major each) in
// one arena allocation.
#pragma once

#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "runtime.hpp"

namespace qtt {

struct Shape {
  Groups g;
  std::vector<std::vector<int64_t>> secs;  // section sizes per dimension
  std::vector<std::vector<Charge>> cvals;  // section charges per dimension
  Charge sel;                              // selection rule

  int rank() const { return int(secs.size()); }
  int64_t sec(int d, int i) const { return secs[d][i]; }
  int nsec(int d) const { return int(secs[d].size()); }
  std::vector<int64_t> sizes() const;
  Charge charge_sum(const BIdx& b) const;
  bool allowed(const BIdx& b) const { return charge_sum(b) == sel; }
  int64_t block_numel(const BIdx& b) const;
  void block_dims(const BIdx& b, int64_t* out) const;
  // member shape_from (sources/btensor.cpp:492-583)
  Shape shape_from(const std::vector<int64_t>& dims) const;
  Shape inverse_cvals() const;
  Shape& neutral_shape_();
  void validate() const;  // structural consistency, throws InvalidArgument
};

Shape shape_product(const std::vector<const Shape*>& parts);  // free shape_from(a, b, ...)

struct Block {
  BIdx idx;
  int64_t off;    // element offset inside the owning storage
  int64_t numel;
};

class BTensor {
 public:
  Context* ctx = nullptr;
  Shape shape;
  std::vector<Block> blocks;  // sorted lexicographically by idx
  BufPtr storage;             // owns (or shares) the block data

  BTensor() = default;
  BTensor(Context* c, Shape s) : ctx(c), shape(std::move(s)) {}

  int rank() const { return shape.rank(); }
  double* ptr(const Block& b) const { return storage->dptr() + b.off; }
  double* base() const { return storage ? storage->dptr() : nullptr; }
  int64_t numel() const;
  const Block* find(const BIdx& idx) const;
  // Allocates storage for the given (sorted, allowed) block list.
  void allocate(const std::vector<BIdx>& idx);
  void check_blocks_allowed() const;  // check_tensor (sources/btensor.cpp:405-467)

  // host transfer (blocks in list order, each row-major)
  void upload(const std::vector<BIdx>& idx, const double* host);
  std::vector<double> download() const;
};

// ---- structural transforms (host bookkeeping; data shared or copied) ------
BTensor reshape(const BTensor& t, const std::vector<int64_t>& index_groups);
BTensor reshape_as(const BTensor& t, const Shape& other, bool overwrite_c_vals = false);
BTensor inverse_cvals(const BTensor& t);
BTensor conj(const BTensor& t);  // real data: inverse charges only
BTensor permute(const BTensor& t, const std::vector<int64_t>& perm);
BTensor clone(const BTensor& t);

// ---- contraction --------------------------------------------------------
struct ContractStats {
  double flops = 0;      // sum over block pairs of 2*M*N*K
  double bytes = 0;      // algorithmic bytes: A blocks + B blocks read, C written
  long out_blocks = 0, pairs = 0, tiles = 0;
};
BTensor tensordot(const BTensor& a, const BTensor& b, const std::vector<int64_t>& da,
                  const std::vector<int64_t>& db, ContractStats* stats = nullptr);
// A contraction planned once and launched any number of times (for graph
// capture): its launch descriptors and temporaries live in arena regions the
// plan owns; out is the result tensor every launch writes.
struct TdPlan {
  struct Launch;
  BTensor out;
  std::vector<BufPtr> keep;
  std::shared_ptr<Launch> launch;
};
TdPlan plan_tensordot(const BTensor& a, const BTensor& b, const std::vector<int64_t>& da,
                      const std::vector<int64_t>& db, ContractStats* stats = nullptr);
// Enqueues the plan's kernels on its context's stream (timed: per-kernel
// profiling events, which must be off under stream capture).
void run_tensordot(const TdPlan& p, bool timed);
// The process-wide table of FP64 GEMM tile variants per contraction shape
// class (filled by timing the variants on first use), as JSON.
void gemm_tuning_save(const std::string& path);
void gemm_tuning_load(const std::string& path);
size_t gemm_tuning_size();
// tensordot written into an existing tensor that has the result's structure
// (same sections, charges, selection rule and block list); every temporary
// comes from the device's static scratch, so nothing is allocated.
void tensordot_into(const BTensor& a, const BTensor& b, const std::vector<int64_t>& da,
                    const std::vector<int64_t>& db, BTensor& out, ContractStats* stats = nullptr);
bool check_product_compat(const Shape& a, const Shape& b, const std::vector<int64_t>& da,
                          const std::vector<int64_t>& db, bool throws);

// ---- elementwise ----------------------------------------------------------
BTensor mul(const BTensor& x, const BTensor& y);   // broadcast mul (btensor.cpp:1631-1689)
BTensor div_scalar_tensor(const BTensor& x, const BTensor& s, double min_abs);  // x / s when |s| >= min_abs
BTensor add(const BTensor& x, const BTensor& y, double alpha);  // flat_map::merge semantics
BTensor axpby_dev(const BTensor& x, const BTensor& a, const BTensor& y, const BTensor& b);  // a*x + b*y
BTensor scale(const BTensor& x, double alpha);
BTensor full_dot(const BTensor& x, const BTensor& y, bool sqrt_out = false);  // rank-0: all dims contracted
BTensor fill_like(Context* ctx, const Shape& s, double value);  // every allowed block filled (ones_like)
double item(const BTensor& scalar);                    // synchronizing read of a rank-0 tensor

// ---- block linear algebra -------------------------------------------------
struct SvdStats {
  double flops = 0;   // sum over sectors 14*m*n^2 + 8*n^3 (m >= n)
  int sectors = 0, sweeps = 0, max_k = 0;
  int unconverged = 0;  // sectors whose Jacobi sweeps hit the sweep cap
};
struct SvdResult {
  BTensor U, d, V;
};
SvdResult svd(const BTensor& t, size_t split, SvdStats* stats = nullptr);
SvdResult svd_truncated(const BTensor& t, size_t split, double tol, size_t min_size, size_t max_size,
                        double pow, SvdStats* stats = nullptr);
struct EighResult {
  BTensor d, U;
};
EighResult eigh(const BTensor& t, size_t split);
SvdResult truncate(SvdResult&& udv, size_t max_size, size_t min_size, double tol, double pow);
EighResult truncate(EighResult&& e_s, size_t max_size, size_t min_size, double tol, double pow);
EighResult eigh_truncated(const BTensor& t, size_t split, double tol, size_t min_size, size_t max_size, double pow);
SvdResult svd_matrix(const BTensor& t, bool some, bool compute_uv);
// Normalises d in place on device: d /= sqrt(sum(d^2)).
void normalize_d(BTensor& d);

// charge utilities used by structural code
std::string fmt_idx(const BIdx& b, int r);

}  // namespace qtt
