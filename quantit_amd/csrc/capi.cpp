// extern "C" boundary of the library: plain pointers and sizes, exceptions
// turned into status codes (include/quantit_amd.h).
#include "../../include/quantit_amd.h"

#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "btensor.hpp"
#include "dmrg.hpp"
#include "kernels.hpp"

// A context is deleted when it has been destroyed AND its last tensor handle
// is gone: tensors reference their context (stream, staging), so a handle
// that outlives qtt_context_destroy keeps the context alive, and every call
// on such a handle except qtt_btensor_destroy fails with
// QTT_INVALID_ARGUMENT.
struct qtt_context {
  qtt::Context ctx;
  std::mutex mu;
  long handles = 0;
  bool closed = false;
  qtt_context(int dev, size_t bytes) : ctx(dev, bytes) { ctx.owner = this; }
};
struct qtt_btensor {
  qtt::BTensor t;
  qtt_context* owner;
};

namespace {

thread_local std::string g_err;
thread_local double g_last_flops = 0.0;

template <class F>
int guard(F&& f) {
  try {
    f();
    return QTT_OK;
  } catch (const qtt::Error& e) {
    g_err = e.what();
    return e.status;
  } catch (const std::bad_alloc& e) {
    g_err = e.what();
    return QTT_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    g_err = e.what();
    return QTT_RUNTIME_ERROR;
  }
}

void need(const void* p, const char* what) {
  if (!p) throw qtt::InvalidArgument(std::string("null pointer: ") + what);
}

qtt_context* owner_of(const qtt::BTensor& t) { return static_cast<qtt_context*>(t.ctx->owner); }

qtt_btensor* wrap(qtt::BTensor&& t) {
  qtt_context* c = owner_of(t);
  {
    std::lock_guard<std::mutex> lk(c->mu);
    ++c->handles;
  }
  return new qtt_btensor{std::move(t), c};
}

// The tensor behind a handle, checked: non-null and its context not destroyed.
const qtt::BTensor& live(const qtt_btensor* h, const char* what) {
  need(h, what);
  std::lock_guard<std::mutex> lk(h->owner->mu);
  if (h->owner->closed) throw qtt::InvalidArgument(std::string(what) + ": its context was destroyed");
  return h->t;
}

// Every tensor of one call must belong to one context (one stream, one arena).
void same_context(const qtt::BTensor& a, const qtt::BTensor& b) {
  if (a.ctx != b.ctx) throw qtt::InvalidArgument("tensors belong to different contexts");
}

void release(qtt_btensor* h) {
  if (!h) return;
  qtt_context* c = h->owner;
  delete h;  // the tensor's storage goes back to the arena first
  bool last = false;
  {
    std::lock_guard<std::mutex> lk(c->mu);
    last = --c->handles == 0 && c->closed;
  }
  if (last) delete c;
}

}  // namespace

extern "C" {

const char* qtt_last_error(void) { return g_err.c_str(); }
int qtt_version(void) { return 100; }

int qtt_context_create(int device, size_t arena_bytes, qtt_context** out) {
  return guard([&] {
    need(out, "out");
    *out = new qtt_context(device, arena_bytes);
  });
}

int qtt_context_destroy(qtt_context* ctx) {
  return guard([&] {
    if (!ctx) return;
    bool now = false;
    {
      std::lock_guard<std::mutex> lk(ctx->mu);
      if (ctx->closed) throw qtt::InvalidArgument("context destroyed twice");
      ctx->closed = true;
      now = ctx->handles == 0;
    }
    if (now) delete ctx;
  });
}

int qtt_context_set_precision(qtt_context* c, int precision) {
  return guard([&] {
    need(c, "ctx");
    if (precision != QTT_PRECISION_FP64 && precision != QTT_PRECISION_SPLIT_FP32)
      throw qtt::InvalidArgument("unknown precision mode");
    c->ctx.precision = qtt::Precision(precision);
  });
}

int qtt_context_set_per_block(qtt_context* c, int enable) {
  return guard([&] {
    need(c, "ctx");
    c->ctx.per_block_launch = enable != 0;
  });
}

int qtt_context_set_overlap(qtt_context* c, int enable) {
  return guard([&] {
    need(c, "ctx");
    c->ctx.overlap = enable != 0;
  });
}

int qtt_context_set_profiling(qtt_context* c, int enable) {
  return guard([&] {
    need(c, "ctx");
    if (c->ctx.profiling && !enable) c->ctx.flush_timings();
    c->ctx.profiling = enable != 0;
  });
}

int qtt_context_sync(qtt_context* c) {
  return guard([&] {
    need(c, "ctx");
    c->ctx.sync();
  });
}

int qtt_context_memory(qtt_context* c, size_t* used, size_t* peak, size_t* capacity) {
  return guard([&] {
    need(c, "ctx");
    if (used) *used = c->ctx.arena().used();
    if (peak) *peak = c->ctx.arena().peak();
    if (capacity) *capacity = c->ctx.arena().capacity();
  });
}

int qtt_context_kernel_times(qtt_context* c, char* buf, size_t buflen, int reset) {
  return guard([&] {
    need(c, "ctx");
    c->ctx.flush_timings();
    std::string s;
    for (auto& kv : c->ctx.kernel_ms) {
      char line[256];
      std::snprintf(line, sizeof(line), "%s %.6f %ld\n", kv.first.c_str(), kv.second.first, kv.second.second);
      s += line;
    }
    if (buf && buflen) {
      std::strncpy(buf, s.c_str(), buflen - 1);
      buf[buflen - 1] = 0;
    }
    if (reset) c->ctx.kernel_ms.clear();
  });
}

int qtt_context_svd_stats(qtt_context* c, int64_t* calls, int64_t* max_sweeps, int64_t* unconverged) {
  return guard([&] {
    need(c, "ctx");
    if (calls) *calls = c->ctx.svd_calls;
    if (max_sweeps) *max_sweeps = c->ctx.svd_max_sweeps;
    if (unconverged) *unconverged = c->ctx.svd_unconverged;
  });
}

int qtt_context_svd_flops(qtt_context* c, double* all, double* wide) {
  return guard([&] {
    need(c, "ctx");
    if (all) *all = c->ctx.svd_model_flops;
    if (wide) *wide = c->ctx.svd_model_flops_wide;
  });
}

int qtt_context_set_lanczos(qtt_context* c, int max_krylov, double tol) {
  return guard([&] {
    need(c, "ctx");
    if (max_krylov != 0 && (max_krylov < 2 || max_krylov > qtt::kern::kMaxKrylov))
      throw qtt::InvalidArgument("max_krylov must be 0 or lie in [2, " + std::to_string(qtt::kern::kMaxKrylov) + "]");
    if (!(tol >= 0.0)) throw qtt::InvalidArgument("tolerance must be non-negative");
    c->ctx.lanczos_krylov = max_krylov;
    c->ctx.lanczos_tol = tol;
  });
}

int qtt_context_set_graphs(qtt_context* c, int enable) {
  return guard([&] {
    need(c, "ctx");
    c->ctx.graphs = enable != 0;
  });
}

int qtt_context_lanczos_stats(qtt_context* c, int64_t* updates, int64_t* vectors, int64_t* graph_replays) {
  return guard([&] {
    need(c, "ctx");
    if (updates) *updates = c->ctx.lanczos_updates;
    if (vectors) *vectors = c->ctx.lanczos_vectors;
    if (graph_replays) *graph_replays = c->ctx.graph_replays;
  });
}

int qtt_btensor_create(qtt_context* c, int rank, const int64_t* nsec, const int64_t* sec_sizes, int ngroups,
                       const uint16_t* group_mod, const int16_t* cvals, const int16_t* sel, int64_t nblocks,
                       const int32_t* block_idx, const double* data, qtt_btensor** out) {
  return guard([&] {
    need(c, "ctx");
    need(out, "out");
    if (rank < 0 || rank > qtt::kMaxRank) throw qtt::InvalidArgument("rank out of range");
    if (ngroups < 1 || ngroups > qtt::kMaxGroups) throw qtt::InvalidArgument("ngroups out of range");
    if (rank) {
      need(nsec, "nsec");
      for (int d = 0; d < rank; ++d)
        if (nsec[d] < 0) throw qtt::InvalidArgument("negative section count");
    }
    qtt::Shape s;
    s.g.n = ngroups;
    for (int i = 0; i < ngroups; ++i) s.g.mod[i] = group_mod ? group_mod[i] : 0;
    int64_t pos = 0;
    for (int d = 0; d < rank; ++d) {
      std::vector<int64_t> sz;
      std::vector<qtt::Charge> cv;
      for (int64_t k = 0; k < nsec[d]; ++k, ++pos) {
        sz.push_back(sec_sizes[pos]);
        qtt::Charge ch{};
        for (int gi = 0; gi < ngroups; ++gi) ch.v[gi] = cvals[pos * ngroups + gi];
        cv.push_back(ch);
      }
      s.secs.push_back(sz);
      s.cvals.push_back(cv);
    }
    for (int gi = 0; gi < ngroups; ++gi) s.sel.v[gi] = sel ? sel[gi] : 0;
    s.validate();
    qtt::BTensor t(&c->ctx, s);
    if (nblocks < 0) throw qtt::InvalidArgument("negative block count");
    if (nblocks && rank) need(block_idx, "block_idx");
    std::vector<qtt::BIdx> idx(static_cast<size_t>(nblocks));
    for (int64_t b = 0; b < nblocks; ++b)
      for (int d = 0; d < rank; ++d) {
        const int32_t k = block_idx[b * rank + d];
        if (k < 0 || k >= s.nsec(d))
          throw qtt::InvalidArgument("block " + std::to_string(b) + ": section index " + std::to_string(k) +
                                     " out of range along dimension " + std::to_string(d));
        idx[b].i[d] = k;
      }
    int64_t n = 0;
    for (auto& b : idx) n += s.block_numel(b);
    if (n) need(data, "data");
    t.upload(idx, data);
    *out = wrap(std::move(t));
  });
}

int qtt_btensor_destroy(qtt_btensor* t) {
  return guard([&] { release(t); });
}

int qtt_btensor_info(const qtt_btensor* h, int* rank, int* ngroups, int64_t* total_sections, int64_t* nblocks,
                     int64_t* numel) {
  return guard([&] {
    const qtt::BTensor& t = live(h, "tensor");
    if (rank) *rank = t.rank();
    if (ngroups) *ngroups = t.shape.g.n;
    int64_t ts = 0;
    for (auto& s : t.shape.secs) ts += int64_t(s.size());
    if (total_sections) *total_sections = ts;
    if (nblocks) *nblocks = int64_t(t.blocks.size());
    if (numel) *numel = t.numel();
  });
}

int qtt_btensor_structure(const qtt_btensor* h, int64_t* nsec, int64_t* sec_sizes, int16_t* cvals, int16_t* sel,
                          uint16_t* group_mod) {
  return guard([&] {
    const auto& s = live(h, "tensor").shape;
    const int ng = s.g.n;
    int64_t pos = 0;
    for (int d = 0; d < s.rank(); ++d) {
      if (nsec) nsec[d] = s.nsec(d);
      for (int k = 0; k < s.nsec(d); ++k, ++pos) {
        if (sec_sizes) sec_sizes[pos] = s.secs[d][k];
        if (cvals)
          for (int gi = 0; gi < ng; ++gi) cvals[pos * ng + gi] = s.cvals[d][k].v[gi];
      }
    }
    if (sel)
      for (int gi = 0; gi < ng; ++gi) sel[gi] = s.sel.v[gi];
    if (group_mod)
      for (int gi = 0; gi < ng; ++gi) group_mod[gi] = s.g.mod[gi];
  });
}

int qtt_btensor_blocks(const qtt_btensor* h, int32_t* block_idx, double* data) {
  return guard([&] {
    const qtt::BTensor& t = live(h, "tensor");
    const int r = t.rank();
    if (block_idx)
      for (size_t b = 0; b < t.blocks.size(); ++b)
        for (int d = 0; d < r; ++d) block_idx[b * r + d] = t.blocks[b].idx.i[d];
    if (data) {
      auto hd = t.download();
      if (!hd.empty()) std::memcpy(data, hd.data(), hd.size() * sizeof(double));
    }
  });
}

int qtt_permute(const qtt_btensor* h, const int64_t* perm, qtt_btensor** out) {
  return guard([&] {
    const qtt::BTensor& t = live(h, "tensor");
    need(out, "out");
    if (t.rank()) need(perm, "perm");
    std::vector<int64_t> p(perm, perm + t.rank());
    *out = wrap(qtt::permute(t, p));
  });
}

int qtt_reshape(const qtt_btensor* h, int ngroups, const int64_t* groups, qtt_btensor** out) {
  return guard([&] {
    const qtt::BTensor& t = live(h, "tensor");
    need(out, "out");
    if (ngroups < 0) throw qtt::InvalidArgument("negative group count");
    if (ngroups) need(groups, "index_groups");
    std::vector<int64_t> g(groups, groups + ngroups);
    *out = wrap(qtt::reshape(t, g));
  });
}

int qtt_conj(const qtt_btensor* h, qtt_btensor** out) {
  return guard([&] {
    const qtt::BTensor& t = live(h, "tensor");
    need(out, "out");
    *out = wrap(qtt::conj(t));
  });
}

int qtt_tensordot(const qtt_btensor* a, const qtt_btensor* b, int ndims, const int64_t* dims_a, const int64_t* dims_b,
                  qtt_btensor** out) {
  return guard([&] {
    const qtt::BTensor& ta = live(a, "a");
    const qtt::BTensor& tb = live(b, "b");
    need(out, "out");
    same_context(ta, tb);
    if (ndims < 0) throw qtt::InvalidArgument("negative number of contracted dimensions");
    if (ndims) {
      need(dims_a, "dims_a");
      need(dims_b, "dims_b");
    }
    std::vector<int64_t> da(dims_a, dims_a + ndims), db(dims_b, dims_b + ndims);
    qtt::ContractStats st;
    *out = wrap(qtt::tensordot(ta, tb, da, db, &st));
    g_last_flops = st.flops;
  });
}

double qtt_last_tensordot_flops(void) { return g_last_flops; }

int qtt_tensordot_into(const qtt_btensor* a, const qtt_btensor* b, int ndims, const int64_t* dims_a,
                       const int64_t* dims_b, qtt_btensor* out) {
  return guard([&] {
    const qtt::BTensor& ta = live(a, "a");
    const qtt::BTensor& tb = live(b, "b");
    live(out, "out");
    same_context(ta, tb);
    same_context(ta, out->t);
    if (ndims < 0) throw qtt::InvalidArgument("negative number of contracted dimensions");
    if (ndims) {
      need(dims_a, "dims_a");
      need(dims_b, "dims_b");
    }
    std::vector<int64_t> da(dims_a, dims_a + ndims), db(dims_b, dims_b + ndims);
    qtt::ContractStats st;
    qtt::tensordot_into(ta, tb, da, db, out->t, &st);
    g_last_flops = st.flops;
  });
}

int qtt_gemm_tuning_save(const char* path) {
  return guard([&] {
    need(path, "path");
    qtt::gemm_tuning_save(path);
  });
}

int qtt_gemm_tuning_load(const char* path) {
  return guard([&] {
    need(path, "path");
    qtt::gemm_tuning_load(path);
  });
}

int qtt_gemm_tuning_size(int64_t* entries) {
  return guard([&] {
    need(entries, "entries");
    *entries = int64_t(qtt::gemm_tuning_size());
  });
}

int qtt_mul(const qtt_btensor* x, const qtt_btensor* y, qtt_btensor** out) {
  return guard([&] {
    const qtt::BTensor& tx = live(x, "x");
    const qtt::BTensor& ty = live(y, "y");
    need(out, "out");
    same_context(tx, ty);
    *out = wrap(qtt::mul(tx, ty));
  });
}

int qtt_add(const qtt_btensor* x, const qtt_btensor* y, double alpha, qtt_btensor** out) {
  return guard([&] {
    const qtt::BTensor& tx = live(x, "x");
    const qtt::BTensor& ty = live(y, "y");
    need(out, "out");
    same_context(tx, ty);
    *out = wrap(qtt::add(tx, ty, alpha));
  });
}

int qtt_full_dot(const qtt_btensor* x, const qtt_btensor* y, double* out) {
  return guard([&] {
    const qtt::BTensor& tx = live(x, "x");
    const qtt::BTensor& ty = live(y, "y");
    need(out, "out");
    same_context(tx, ty);
    *out = qtt::item(qtt::full_dot(tx, ty));
  });
}

int qtt_svd(const qtt_btensor* t, size_t split, qtt_btensor** U, qtt_btensor** d, qtt_btensor** V) {
  return guard([&] {
    const qtt::BTensor& tt = live(t, "tensor");
    need(U, "U");
    need(d, "d");
    need(V, "V");
    auto r = qtt::svd(tt, split);
    *U = wrap(std::move(r.U));
    *d = wrap(std::move(r.d));
    *V = wrap(std::move(r.V));
  });
}

int qtt_svd_truncated(const qtt_btensor* t, size_t split, double tol, size_t min_size, size_t max_size, double pow,
                      qtt_btensor** U, qtt_btensor** d, qtt_btensor** V) {
  return guard([&] {
    const qtt::BTensor& tt = live(t, "tensor");
    need(U, "U");
    need(d, "d");
    need(V, "V");
    auto r = qtt::svd_truncated(tt, split, tol, min_size, max_size, pow);
    *U = wrap(std::move(r.U));
    *d = wrap(std::move(r.d));
    *V = wrap(std::move(r.V));
  });
}

int qtt_eigh(const qtt_btensor* t, size_t split, qtt_btensor** d, qtt_btensor** U) {
  return guard([&] {
    const qtt::BTensor& tt = live(t, "tensor");
    need(U, "U");
    need(d, "d");
    auto r = qtt::eigh(tt, split);
    *d = wrap(std::move(r.d));
    *U = wrap(std::move(r.U));
  });
}

int qtt_truncate(const qtt_btensor* U, const qtt_btensor* d, const qtt_btensor* V, size_t max_size, size_t min_size,
                 double tol, double pow, qtt_btensor** U_out, qtt_btensor** d_out, qtt_btensor** V_out) {
  return guard([&] {
    const qtt::BTensor& tu = live(U, "U");
    const qtt::BTensor& td = live(d, "d");
    const qtt::BTensor& tv = live(V, "V");
    need(U_out, "U_out");
    need(d_out, "d_out");
    need(V_out, "V_out");
    same_context(tu, td);
    same_context(tu, tv);
    // the shapes an svd returns: d rank 1, and the last legs of U and V carry
    // d's sections (the truncation resizes them section by section)
    if (td.rank() != 1) throw qtt::InvalidArgument("truncate: d must have rank 1");
    if (tu.rank() < 1 || tv.rank() < 1) throw qtt::InvalidArgument("truncate: U and V need at least one dimension");
    const int nd = td.shape.nsec(0);
    if (tu.shape.nsec(tu.rank() - 1) != nd || tv.shape.nsec(tv.rank() - 1) != nd)
      throw qtt::InvalidArgument("truncate: the last dimensions of U and V must have the sections of d");
    for (auto& b : td.blocks)
      if (tu.shape.secs[tu.rank() - 1][b.idx.i[0]] != td.shape.secs[0][b.idx.i[0]] ||
          tv.shape.secs[tv.rank() - 1][b.idx.i[0]] != td.shape.secs[0][b.idx.i[0]])
        throw qtt::InvalidArgument("truncate: section sizes of U, d and V differ");
    qtt::SvdResult in{tu, td, tv};
    auto r = qtt::truncate(std::move(in), max_size, min_size, tol, pow);
    *U_out = wrap(std::move(r.U));
    *d_out = wrap(std::move(r.d));
    *V_out = wrap(std::move(r.V));
  });
}

int qtt_eigh_truncated(const qtt_btensor* t, size_t split, double tol, size_t min_size, size_t max_size, double pow,
                       qtt_btensor** d, qtt_btensor** U) {
  return guard([&] {
    const qtt::BTensor& tt = live(t, "tensor");
    need(U, "U");
    need(d, "d");
    auto r = qtt::eigh_truncated(tt, split, tol, min_size, max_size, pow);
    *d = wrap(std::move(r.d));
    *U = wrap(std::move(r.U));
  });
}

int qtt_truncate_eigh(const qtt_btensor* e, const qtt_btensor* S, size_t max_size, size_t min_size, double tol,
                      double pow, qtt_btensor** e_out, qtt_btensor** S_out) {
  return guard([&] {
    const qtt::BTensor& te = live(e, "e");
    const qtt::BTensor& ts = live(S, "S");
    need(e_out, "e_out");
    need(S_out, "S_out");
    same_context(te, ts);
    if (te.rank() != 1) throw qtt::InvalidArgument("truncate: e must have rank 1");
    if (ts.rank() < 1) throw qtt::InvalidArgument("truncate: S needs at least one dimension");
    const int nd = te.shape.nsec(0);
    if (ts.shape.nsec(ts.rank() - 1) != nd)
      throw qtt::InvalidArgument("truncate: the last dimension of S must have the sections of e");
    for (auto& b : te.blocks)
      if (ts.shape.secs[ts.rank() - 1][b.idx.i[0]] != te.shape.secs[0][b.idx.i[0]])
        throw qtt::InvalidArgument("truncate: section sizes of e and S differ");
    qtt::EighResult in{te, ts};
    auto r = qtt::truncate(std::move(in), max_size, min_size, tol, pow);
    *e_out = wrap(std::move(r.d));
    *S_out = wrap(std::move(r.U));
  });
}

int qtt_svd_matrix(const qtt_btensor* t, int some, int compute_uv, qtt_btensor** U, qtt_btensor** d,
                   qtt_btensor** V) {
  return guard([&] {
    const qtt::BTensor& tt = live(t, "tensor");
    need(U, "U");
    need(d, "d");
    need(V, "V");
    if (tt.rank() != 2) throw qtt::InvalidArgument("svd(tensor, some, compute_uv): the tensor must have rank 2");
    auto r = qtt::svd_matrix(tt, some != 0, compute_uv != 0);
    *U = wrap(std::move(r.U));
    *d = wrap(std::move(r.d));
    *V = wrap(std::move(r.V));
  });
}

namespace {

// Tensors of an MPS/MPO list, all of one context.
std::vector<qtt::BTensor> gather(int L, qtt_btensor* const* hs, const char* what, const qtt::Context* ctx) {
  need(hs, what);
  std::vector<qtt::BTensor> v;
  for (int i = 0; i < L; ++i) {
    const qtt::BTensor& t = live(hs[i], what);
    if (ctx && t.ctx != ctx) throw qtt::InvalidArgument("tensors belong to different contexts");
    v.push_back(t);
  }
  return v;
}

void check_length(int L) {
  if (L < 1) throw qtt::InvalidArgument("an MPS needs at least one site");
}

// Replaces the handles of a list in place (the old ones are released).
void replace(int L, qtt_btensor** hs, std::vector<qtt::BTensor>& v) {
  for (int i = 0; i < L; ++i) {
    qtt_btensor* n = wrap(std::move(v[i]));
    release(hs[i]);
    hs[i] = n;
  }
}

}  // namespace

int qtt_dmrg(int L, qtt_btensor* const* mpo, qtt_btensor** mps, const qtt_dmrg_options* options, double* energy,
             int max_log, double* sweep_energies, double* sweep_ms, int64_t* max_bond, int* n_sweeps,
             double* contract_flops, double* svd_flops) {
  return qtt_dmrg_from_center(L, mpo, mps, 0, options, energy, max_log, sweep_energies, sweep_ms, max_bond, n_sweeps,
                              contract_flops, svd_flops);
}

int qtt_dmrg_from_center(int L, qtt_btensor* const* mpo, qtt_btensor** mps, int oc, const qtt_dmrg_options* options,
                         double* energy, int max_log, double* sweep_energies, double* sweep_ms, int64_t* max_bond,
                         int* n_sweeps, double* contract_flops, double* svd_flops) {
  return guard([&] {
    need(options, "options");
    check_length(L);
    qtt::DmrgState st;
    st.oc = oc;
    need(mps, "mps");
    st.mps = gather(L, mps, "mps", nullptr);
    st.mpo = gather(L, mpo, "mpo", st.mps[0].ctx);
    for (auto& t : st.mps)
      if (t.ctx != st.mps[0].ctx) throw qtt::InvalidArgument("tensors belong to different contexts");
    qtt::DmrgOptions opt;
    opt.cutoff = options->cutoff;
    opt.convergence_criterion = options->convergence_criterion;
    opt.maximum_bond = size_t(options->maximum_bond);
    opt.minimum_bond = size_t(options->minimum_bond);
    opt.maximum_iterations = size_t(options->maximum_iterations);
    qtt::DmrgLog log;
    const double E = qtt::dmrg_run(st, opt, &log);
    replace(L, mps, st.mps);
    if (energy) *energy = E;
    const int ns = int(log.energies.size());
    if (n_sweeps) *n_sweeps = ns;
    for (int i = 0; i < ns && i < max_log; ++i) {
      if (sweep_energies) sweep_energies[i] = log.energies[i];
      if (sweep_ms) sweep_ms[i] = log.sweep_ms[i];
      if (max_bond) max_bond[i] = log.max_bond[i];
    }
    if (contract_flops) *contract_flops = log.contract_flops;
    if (svd_flops) *svd_flops = log.svd_flops;
  });
}

int qtt_dmrg_logged(int L, qtt_btensor* const* mpo, qtt_btensor** mps, int oc, const qtt_dmrg_options* options,
                    const qtt_dmrg_logger* logger, double* energy) {
  return guard([&] {
    need(options, "options");
    need(logger, "logger");
    check_length(L);
    qtt::DmrgState st;
    st.oc = oc;
    need(mps, "mps");
    st.mps = gather(L, mps, "mps", nullptr);
    st.mpo = gather(L, mpo, "mpo", st.mps[0].ctx);
    for (auto& t : st.mps)
      if (t.ctx != st.mps[0].ctx) throw qtt::InvalidArgument("tensors belong to different contexts");
    qtt::DmrgOptions opt;
    opt.cutoff = options->cutoff;
    opt.convergence_criterion = options->convergence_criterion;
    opt.maximum_bond = size_t(options->maximum_bond);
    opt.minimum_bond = size_t(options->minimum_bond);
    opt.maximum_iterations = size_t(options->maximum_iterations);
    // the state is lent to the callback as handles valid for the call only
    auto lend = [&](void (*fn)(void*, size_t, double, int, qtt_btensor* const*), size_t it, double E,
                    const std::vector<qtt::BTensor>& state) {
      std::vector<qtt_btensor*> hs;
      hs.reserve(state.size());
      for (const auto& t : state) hs.push_back(wrap(qtt::BTensor(t)));
      fn(logger->user, it, E, int(hs.size()), hs.data());
      for (auto* h : hs) release(h);
    };
    qtt::DmrgHooks hooks;
    if (logger->init) hooks.init = [&] { logger->init(logger->user, options); };
    if (logger->it_log_all)
      hooks.it_log_all = [&](size_t it, double E, const std::vector<qtt::BTensor>& s) {
        lend(logger->it_log_all, it, E, s);
      };
    if (logger->end_log_all)
      hooks.end_log_all = [&](size_t it, double E, const std::vector<qtt::BTensor>& s) {
        lend(logger->end_log_all, it, E, s);
      };
    qtt::DmrgLog log;
    const double E = qtt::dmrg_run(st, opt, &log, &hooks);
    replace(L, mps, st.mps);
    if (energy) *energy = E;
  });
}

int qtt_hamil2site_times_state(const qtt_btensor* theta, const qtt_btensor* W12, const qtt_btensor* L,
                               const qtt_btensor* R, qtt_btensor** out) {
  return guard([&] {
    const qtt::BTensor& t = live(theta, "theta");
    const qtt::BTensor& w = live(W12, "W12");
    const qtt::BTensor& l = live(L, "L");
    const qtt::BTensor& r = live(R, "R");
    need(out, "out");
    same_context(t, w);
    same_context(t, l);
    same_context(t, r);
    *out = wrap(qtt::hamil2site_times_state(t, w, l, r));
  });
}

int qtt_one_step_lanczos(const qtt_btensor* theta, const qtt_btensor* W12, const qtt_btensor* L,
                         const qtt_btensor* R, qtt_btensor** Hpsi, double* a0, double* a1, double* b) {
  return guard([&] {
    const qtt::BTensor& t = live(theta, "theta");
    const qtt::BTensor& w = live(W12, "W12");
    const qtt::BTensor& l = live(L, "L");
    const qtt::BTensor& r = live(R, "R");
    need(Hpsi, "Hpsi");
    same_context(t, w);
    same_context(t, l);
    same_context(t, r);
    qtt::LanczosStep st = qtt::one_step_lanczos(t, w, l, r);
    if (a0) *a0 = qtt::item(st.a0);
    if (a1) *a1 = qtt::item(st.a1);
    if (b) *b = qtt::item(st.b);
    *Hpsi = wrap(std::move(st.Hpsi));
  });
}

int qtt_eig2x2(qtt_context* c, double a0, double a1, double b, double* E, double* o, double* n) {
  return guard([&] {
    need(c, "ctx");
    qtt::Shape s0;
    s0.sel = qtt::q_neutral();
    qtt::BTensor x0 = qtt::fill_like(&c->ctx, s0, a0), x1 = qtt::fill_like(&c->ctx, s0, a1),
                 xb = qtt::fill_like(&c->ctx, s0, b);
    qtt::Eig2x2 r = qtt::eig2x2_dev(x0, x1, xb);
    int32_t flag = 0;
    c->ctx.read_small(&flag, r.flag->cptr(), sizeof(int32_t));
    if (flag) throw qtt::LogicError("nan found in output tensor");
    if (E) *E = qtt::item(r.E);
    if (o) *o = qtt::item(r.o);
    if (n) *n = qtt::item(r.n);
  });
}

int qtt_contract_norm(int L, qtt_btensor* const* a, qtt_btensor* const* b, double* out) {
  return guard([&] {
    need(out, "out");
    check_length(L);
    auto va = gather(L, a, "a", nullptr);
    auto vb = gather(L, b, "b", va[0].ctx);
    for (auto& t : va)
      if (t.ctx != va[0].ctx) throw qtt::InvalidArgument("tensors belong to different contexts");
    *out = qtt::contract_norm(va, vb);
  });
}

int qtt_contract_energy(int L, qtt_btensor* const* a, qtt_btensor* const* b, qtt_btensor* const* mpo, double* out) {
  return guard([&] {
    need(out, "out");
    check_length(L);
    auto va = gather(L, a, "a", nullptr);
    auto vb = gather(L, b, "b", va[0].ctx);
    auto vm = gather(L, mpo, "mpo", va[0].ctx);
    for (auto& t : va)
      if (t.ctx != va[0].ctx) throw qtt::InvalidArgument("tensors belong to different contexts");
    *out = qtt::contract_energy(va, vb, vm);
  });
}

int qtt_mps_move_oc(int L, qtt_btensor** mps, int oc, int target) {
  return guard([&] {
    check_length(L);
    auto v = gather(L, mps, "mps", nullptr);
    for (auto& t : v)
      if (t.ctx != v[0].ctx) throw qtt::InvalidArgument("tensors belong to different contexts");
    qtt::mps_move_oc(v, oc, target);
    replace(L, mps, v);
  });
}

int qtt_mps_prepare(int L, qtt_btensor** mps) {
  return guard([&] {
    check_length(L);
    auto v = gather(L, mps, "mps", nullptr);
    for (auto& t : v)
      if (t.ctx != v[0].ctx) throw qtt::InvalidArgument("tensors belong to different contexts");
    qtt::mps_prepare(v);
    replace(L, mps, v);
  });
}

}  // extern "C"
