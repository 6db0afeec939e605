// extern "C" boundary of the library: plain pointers and sizes, exceptions
// turned into status codes (include/quantit_amd.h).
#include "../../include/quantit_amd.h"

#include <cstdio>
#include <cstring>
#include <string>

#include "btensor.hpp"
#include "dmrg.hpp"

struct qtt_context {
  qtt::Context ctx;
  qtt_context(int dev, size_t bytes) : ctx(dev, bytes) {}
};
struct qtt_btensor {
  qtt::BTensor t;
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

qtt_btensor* wrap(qtt::BTensor&& t) { return new qtt_btensor{std::move(t)}; }

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
  return guard([&] { delete ctx; });
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

int qtt_btensor_create(qtt_context* c, int rank, const int64_t* nsec, const int64_t* sec_sizes, int ngroups,
                       const uint16_t* group_mod, const int16_t* cvals, const int16_t* sel, int64_t nblocks,
                       const int32_t* block_idx, const double* data, qtt_btensor** out) {
  return guard([&] {
    need(c, "ctx");
    need(out, "out");
    if (rank < 0 || rank > qtt::kMaxRank) throw qtt::InvalidArgument("rank out of range");
    if (ngroups < 1 || ngroups > qtt::kMaxGroups) throw qtt::InvalidArgument("ngroups out of range");
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
    std::vector<qtt::BIdx> idx(size_t(std::max<int64_t>(nblocks, 0)));
    for (int64_t b = 0; b < nblocks; ++b)
      for (int d = 0; d < rank; ++d) idx[b].i[d] = block_idx[b * rank + d];
    int64_t n = 0;
    for (auto& b : idx) n += s.block_numel(b);
    if (n) need(data, "data");
    t.upload(idx, data);
    *out = wrap(std::move(t));
  });
}

int qtt_btensor_destroy(qtt_btensor* t) {
  return guard([&] { delete t; });
}

int qtt_btensor_info(const qtt_btensor* t, int* rank, int* ngroups, int64_t* total_sections, int64_t* nblocks,
                     int64_t* numel) {
  return guard([&] {
    need(t, "tensor");
    if (rank) *rank = t->t.rank();
    if (ngroups) *ngroups = t->t.shape.g.n;
    int64_t ts = 0;
    for (auto& s : t->t.shape.secs) ts += int64_t(s.size());
    if (total_sections) *total_sections = ts;
    if (nblocks) *nblocks = int64_t(t->t.blocks.size());
    if (numel) *numel = t->t.numel();
  });
}

int qtt_btensor_structure(const qtt_btensor* t, int64_t* nsec, int64_t* sec_sizes, int16_t* cvals, int16_t* sel,
                          uint16_t* group_mod) {
  return guard([&] {
    need(t, "tensor");
    const auto& s = t->t.shape;
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

int qtt_btensor_blocks(const qtt_btensor* t, int32_t* block_idx, double* data) {
  return guard([&] {
    need(t, "tensor");
    const int r = t->t.rank();
    if (block_idx)
      for (size_t b = 0; b < t->t.blocks.size(); ++b)
        for (int d = 0; d < r; ++d) block_idx[b * r + d] = t->t.blocks[b].idx.i[d];
    if (data) {
      auto h = t->t.download();
      if (!h.empty()) std::memcpy(data, h.data(), h.size() * sizeof(double));
    }
  });
}

int qtt_permute(const qtt_btensor* t, const int64_t* perm, qtt_btensor** out) {
  return guard([&] {
    need(t, "tensor");
    need(out, "out");
    std::vector<int64_t> p(perm, perm + t->t.rank());
    *out = wrap(qtt::permute(t->t, p));
  });
}

int qtt_reshape(const qtt_btensor* t, int ngroups, const int64_t* groups, qtt_btensor** out) {
  return guard([&] {
    need(t, "tensor");
    need(out, "out");
    std::vector<int64_t> g(groups, groups + ngroups);
    *out = wrap(qtt::reshape(t->t, g));
  });
}

int qtt_conj(const qtt_btensor* t, qtt_btensor** out) {
  return guard([&] {
    need(t, "tensor");
    need(out, "out");
    *out = wrap(qtt::conj(t->t));
  });
}

int qtt_tensordot(const qtt_btensor* a, const qtt_btensor* b, int ndims, const int64_t* dims_a, const int64_t* dims_b,
                  qtt_btensor** out) {
  return guard([&] {
    need(a, "a");
    need(b, "b");
    need(out, "out");
    if (a->t.ctx != b->t.ctx) throw qtt::InvalidArgument("tensors belong to different contexts");
    std::vector<int64_t> da(dims_a, dims_a + ndims), db(dims_b, dims_b + ndims);
    qtt::ContractStats st;
    *out = wrap(qtt::tensordot(a->t, b->t, da, db, &st));
    g_last_flops = st.flops;
  });
}

double qtt_last_tensordot_flops(void) { return g_last_flops; }

int qtt_mul(const qtt_btensor* x, const qtt_btensor* y, qtt_btensor** out) {
  return guard([&] {
    need(x, "x");
    need(y, "y");
    need(out, "out");
    *out = wrap(qtt::mul(x->t, y->t));
  });
}

int qtt_add(const qtt_btensor* x, const qtt_btensor* y, double alpha, qtt_btensor** out) {
  return guard([&] {
    need(x, "x");
    need(y, "y");
    need(out, "out");
    *out = wrap(qtt::add(x->t, y->t, alpha));
  });
}

int qtt_full_dot(const qtt_btensor* x, const qtt_btensor* y, double* out) {
  return guard([&] {
    need(x, "x");
    need(y, "y");
    need(out, "out");
    *out = qtt::item(qtt::full_dot(x->t, y->t));
  });
}

int qtt_svd(const qtt_btensor* t, size_t split, qtt_btensor** U, qtt_btensor** d, qtt_btensor** V) {
  return guard([&] {
    need(t, "tensor");
    need(U, "U");
    need(d, "d");
    need(V, "V");
    auto r = qtt::svd(t->t, split);
    *U = wrap(std::move(r.U));
    *d = wrap(std::move(r.d));
    *V = wrap(std::move(r.V));
  });
}

int qtt_svd_truncated(const qtt_btensor* t, size_t split, double tol, size_t min_size, size_t max_size, double pow,
                      qtt_btensor** U, qtt_btensor** d, qtt_btensor** V) {
  return guard([&] {
    need(t, "tensor");
    need(U, "U");
    need(d, "d");
    need(V, "V");
    auto r = qtt::svd_truncated(t->t, split, tol, min_size, max_size, pow);
    *U = wrap(std::move(r.U));
    *d = wrap(std::move(r.d));
    *V = wrap(std::move(r.V));
  });
}

int qtt_eigh(const qtt_btensor* t, size_t split, qtt_btensor** d, qtt_btensor** U) {
  return guard([&] {
    need(t, "tensor");
    need(U, "U");
    need(d, "d");
    auto r = qtt::eigh(t->t, split);
    *d = wrap(std::move(r.d));
    *U = wrap(std::move(r.U));
  });
}

int qtt_truncate(const qtt_btensor* U, const qtt_btensor* d, const qtt_btensor* V, size_t max_size, size_t min_size,
                 double tol, double pow, qtt_btensor** U_out, qtt_btensor** d_out, qtt_btensor** V_out) {
  return guard([&] {
    need(U, "U");
    need(d, "d");
    need(V, "V");
    need(U_out, "U_out");
    need(d_out, "d_out");
    need(V_out, "V_out");
    qtt::SvdResult in{U->t, d->t, V->t};
    auto r = qtt::truncate(std::move(in), max_size, min_size, tol, pow);
    *U_out = wrap(std::move(r.U));
    *d_out = wrap(std::move(r.d));
    *V_out = wrap(std::move(r.V));
  });
}

int qtt_dmrg(int L, qtt_btensor* const* mpo, qtt_btensor** mps, const qtt_dmrg_options* options, double* energy,
             int max_log, double* sweep_energies, double* sweep_ms, int64_t* max_bond, int* n_sweeps,
             double* contract_flops, double* svd_flops) {
  return guard([&] {
    need(mpo, "mpo");
    need(mps, "mps");
    need(options, "options");
    qtt::DmrgState st;
    for (int i = 0; i < L; ++i) {
      need(mpo[i], "mpo[i]");
      need(mps[i], "mps[i]");
      st.mpo.push_back(mpo[i]->t);
      st.mps.push_back(mps[i]->t);
    }
    qtt::DmrgOptions opt;
    opt.cutoff = options->cutoff;
    opt.convergence_criterion = options->convergence_criterion;
    opt.maximum_bond = size_t(options->maximum_bond);
    opt.minimum_bond = size_t(options->minimum_bond);
    opt.maximum_iterations = size_t(options->maximum_iterations);
    qtt::DmrgLog log;
    const double E = qtt::dmrg_run(st, opt, &log);
    for (int i = 0; i < L; ++i) {
      delete mps[i];
      mps[i] = wrap(std::move(st.mps[i]));
    }
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

int qtt_contract_norm(int L, qtt_btensor* const* a, qtt_btensor* const* b, double* out) {
  return guard([&] {
    need(a, "a");
    need(b, "b");
    need(out, "out");
    std::vector<qtt::BTensor> va, vb;
    for (int i = 0; i < L; ++i) {
      va.push_back(a[i]->t);
      vb.push_back(b[i]->t);
    }
    *out = qtt::contract_norm(va, vb);
  });
}

int qtt_contract_energy(int L, qtt_btensor* const* a, qtt_btensor* const* b, qtt_btensor* const* mpo, double* out) {
  return guard([&] {
    need(a, "a");
    need(b, "b");
    need(mpo, "mpo");
    need(out, "out");
    std::vector<qtt::BTensor> va, vb, vm;
    for (int i = 0; i < L; ++i) {
      va.push_back(a[i]->t);
      vb.push_back(b[i]->t);
      vm.push_back(mpo[i]->t);
    }
    *out = qtt::contract_energy(va, vb, vm);
  });
}

int qtt_mps_move_oc(int L, qtt_btensor** mps, int oc, int target) {
  return guard([&] {
    need(mps, "mps");
    std::vector<qtt::BTensor> v;
    for (int i = 0; i < L; ++i) {
      need(mps[i], "mps[i]");
      v.push_back(mps[i]->t);
    }
    qtt::mps_move_oc(v, oc, target);
    for (int i = 0; i < L; ++i) {
      delete mps[i];
      mps[i] = wrap(std::move(v[i]));
    }
  });
}

int qtt_mps_prepare(int L, qtt_btensor** mps) {
  return guard([&] {
    need(mps, "mps");
    std::vector<qtt::BTensor> v;
    for (int i = 0; i < L; ++i) {
      need(mps[i], "mps[i]");
      v.push_back(mps[i]->t);
    }
    qtt::mps_prepare(v);
    for (int i = 0; i < L; ++i) {
      delete mps[i];
      mps[i] = wrap(std::move(v[i]));
    }
  });
}

}  // extern "C"
