// k-step Lanczos with full FP64 reorthogonalisation for the two-site update:
// the production eigensolver beside the reference's one-step form
// (sources/dmrg.cpp:584-608, which it reduces to for two Krylov vectors).
//
// The Krylov vectors live in one flat layout: the blocks of the two-site
// state, in block order, back to back; a vector is a row of V.  The layout
// starts from theta's blocks and grows when a matvec output has a block
// outside it (earlier rows get zeros there).  Per step: one matvec (three
// tensordots), its output scattered into w, two classical Gram-Schmidt passes
// against all rows (a skinny product and its transpose each), the norm, and
// the tridiagonal eigensolve; the host reads three numbers (eigenvalue,
// residual, beta) to decide whether to stop.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "dmrg.hpp"
#include "kernels.hpp"

namespace qtt {

BTensor matvec_labelled(const BTensor& x, const BTensor& W, const BTensor& Lenv, const BTensor& Renv,
                        ContractStats* cs);

namespace {

void add_chunks(std::vector<kern::ChunkTask>& tasks, int desc, int64_t numel) {
  const int64_t nch = (numel + kern::kChunk - 1) / kern::kChunk;
  for (int64_t c = 0; c < nch; ++c) tasks.push_back({desc, int32_t(c)});
}

struct Layout {
  int rank = 0;
  std::vector<BIdx> idx;  // sorted
  std::vector<int64_t> off, numel;
  int64_t n = 0;
  int find(const BIdx& b) const {
    auto it = std::lower_bound(idx.begin(), idx.end(), b,
                               [this](const BIdx& x, const BIdx& y) { return bidx_less(x, y, rank); });
    if (it != idx.end() && bidx_eq(*it, b, rank)) return int(it - idx.begin());
    return -1;
  }
  void add(const BIdx& b, int64_t ne) {
    idx.push_back(b);
    numel.push_back(ne);
  }
  void finish() {
    std::vector<int> ord(idx.size());
    for (size_t i = 0; i < ord.size(); ++i) ord[i] = int(i);
    std::sort(ord.begin(), ord.end(), [this](int x, int y) { return bidx_less(idx[x], idx[y], rank); });
    std::vector<BIdx> i2;
    std::vector<int64_t> n2;
    for (int o : ord) {
      i2.push_back(idx[o]);
      n2.push_back(numel[o]);
    }
    idx.swap(i2);
    numel.swap(n2);
    off.assign(idx.size(), 0);
    n = 0;
    for (size_t i = 0; i < idx.size(); ++i) {
      off[i] = n;
      n += numel[i];
    }
  }
};

void launch_copies(Context* ctx, double* ab, const std::vector<kern::Copy2D>& cp,
                   const std::vector<kern::ChunkTask>& tasks, const char* name) {
  if (tasks.empty()) return;
  KernelTimer kt(*ctx, name);
  auto [dc, dt] = ctx->upload_pack(cp, tasks);
  kern::copy2d(ctx->stream(), ab, dc, dt, int(tasks.size()));
}

kern::Copy2D flat_copy(int64_t src, int64_t dst, int64_t ne) {
  kern::Copy2D c{};
  c.src_off = src;  // < 0: fill with alpha
  c.dst_off = dst;
  c.rows = 1;
  c.cols = int32_t(ne);
  c.src_cs = 1;
  c.dst_cs = 1;
  c.alpha = src < 0 ? 0.0 : 1.0;
  return c;
}

// The matvec L.x.W12.R of one update as a HIP graph: three planned
// contractions captured once and replayed for every Krylov vector whose
// layout is the one it was built for.  x is copied into the graph's own input
// tensor before each replay; the result lands in the last plan's output.
class MatvecGraph {
 public:
  MatvecGraph() = default;
  MatvecGraph(const MatvecGraph&) = delete;
  MatvecGraph& operator=(const MatvecGraph&) = delete;
  ~MatvecGraph() { reset(); }
  void reset() {
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    exec_ = nullptr;
    graph_ = nullptr;
    p1_ = p2_ = p3_ = TdPlan();
    x_ = BTensor();
    n_ = -1;
  }
  bool ready(int64_t layout_n) const { return exec_ && layout_n == n_; }
  // Plans the three contractions for inputs of `layout` (the blocks of x, in
  // order) and captures their kernels.
  void build(const BTensor& shape_of_x, const Layout& lay, const BTensor& W, const BTensor& L, const BTensor& R) {
    reset();
    Context* ctx = shape_of_x.ctx;
    x_ = BTensor(ctx, shape_of_x.shape);
    x_.allocate(lay.idx);
    ContractStats cs;
    p1_ = plan_tensordot(L, x_, {0}, {0}, &cs);
    p2_ = plan_tensordot(p1_.out, W, {0, 2, 3}, {0, 4, 5}, &cs);
    p3_ = plan_tensordot(p2_.out, R, {1, 4}, {0, 1}, &cs);
    flops_ = cs.flops;
    QTT_HIP(hipStreamBeginCapture(ctx->stream(), hipStreamCaptureModeThreadLocal));
    run_tensordot(p1_, false);
    run_tensordot(p2_, false);
    run_tensordot(p3_, false);
    QTT_HIP(hipStreamEndCapture(ctx->stream(), &graph_));
    QTT_HIP(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
    n_ = lay.n;
  }
  // x (row `row` of V, pitch n) -> graph input, then the replay.
  void launch(Context* ctx, int64_t v_off, ContractStats* cs) {
    double* ab = reinterpret_cast<double*>(ctx->arena().base());
    KernelTimer kt(*ctx, "lanczos_matvec_graph");
    kern::krylov_axpy(ctx->stream(), ab, 0, 0, 0, 0, 0.0, v_off, x_.base() - ab, n_, -1);
    QTT_HIP(hipGraphLaunch(exec_, ctx->stream()));
    if (cs) cs->flops += flops_;
    ++replays;
  }
  const BTensor& out() const { return p3_.out; }
  int64_t replays = 0;

 private:
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
  TdPlan p1_, p2_, p3_;
  BTensor x_;
  int64_t n_ = -1;
  double flops_ = 0;
};

}  // namespace

LanczosResult lanczos_ground(const BTensor& theta, const BTensor& W, const BTensor& Lenv, const BTensor& Renv,
                             int max_krylov, double tol, ContractStats* cs) {
  Context* ctx = theta.ctx;
  double* ab = reinterpret_cast<double*>(ctx->arena().base());
  const int K = max_krylov;
  if (K < 2 || K > kern::kMaxKrylov)
    throw InvalidArgument("Krylov dimension must lie in [2, " + std::to_string(kern::kMaxKrylov) + "]");
  Layout lay;
  lay.rank = theta.rank();
  for (auto& b : theta.blocks)
    if (b.numel) lay.add(b.idx, b.numel);
  lay.finish();
  std::vector<char> present(lay.idx.size(), 1);  // blocks some Krylov vector so far has
  if (lay.n == 0) throw DomainError("two-site state without elements");

  // scalars: c1[K], c2[K], nrm2, T[2K], out[3 + K], then the NaN flag (an
  // int32 in a double slot), so one read per step brings both back
  BufPtr sc = ctx->alloc(size_t(6 * K + 8) * sizeof(double));
  const int64_t s0 = sc->dptr() - ab;
  const int64_t c1 = s0, c2 = s0 + K, nrm2 = s0 + 2 * K, t_off = s0 + 2 * K + 1, out = s0 + 4 * K + 1;
  int32_t* d_flag = reinterpret_cast<int32_t*>(ab + out + 3 + K);
  kern::fill_i32(ctx->stream(), d_flag, 1, 0);
  BufPtr V, w, part;
  auto alloc_vectors = [&](int64_t n) {
    V = ctx->alloc(size_t(K) * size_t(n) * sizeof(double));
    w = ctx->alloc(size_t(n) * sizeof(double));
    part = ctx->alloc(size_t(K) * size_t((n + kern::kKrylovChunk - 1) / kern::kKrylovChunk + 1) * sizeof(double));
  };
  alloc_vectors(lay.n);
  auto voff = [&](int row) { return V->dptr() - ab + int64_t(row) * lay.n; };
  auto woff = [&] { return w->dptr() - ab; };
  auto poff = [&] { return part->dptr() - ab; };

  // v0 = theta / |theta|
  {
    std::vector<kern::Copy2D> cp;
    std::vector<kern::ChunkTask> tasks;
    for (auto& b : theta.blocks) {
      if (!b.numel) continue;
      const int k = lay.find(b.idx);
      add_chunks(tasks, int(cp.size()), b.numel);
      cp.push_back(flat_copy(theta.ptr(b) - ab, voff(0) + lay.off[k], b.numel));
    }
    launch_copies(ctx, ab, cp, tasks, "lanczos_scatter");
    KernelTimer kt(*ctx, "lanczos_vec");
    kern::krylov_dot(ctx->stream(), ab, voff(0), lay.n, 1, voff(0), lay.n, poff(), nrm2);
    kern::krylov_axpy(ctx->stream(), ab, voff(0), lay.n, 0, nrm2, 0.0, voff(0), voff(0), lay.n, nrm2);
  }

  LanczosResult res;
  int m = 0;
  double host_out[3] = {0, 0, 0};
  MatvecGraph graph;
  for (int j = 0; j < K; ++j) {
    // H v_j: eagerly through a view of row j, or (from the second vector on,
    // when graphs are enabled) by replaying the captured matvec of the
    // current layout
    BTensor o;
    if (ctx->graphs && j >= 1) {
      if (!graph.ready(lay.n)) {
        BTensor x(ctx, theta.shape);
        graph.build(x, lay, W, Lenv, Renv);
      }
      graph.launch(ctx, voff(j), cs);
      o = graph.out();
    } else {
      BTensor x(ctx, theta.shape);
      x.storage = V;
      for (size_t b = 0; b < lay.idx.size(); ++b)
        x.blocks.push_back({lay.idx[b], int64_t(j) * lay.n + lay.off[b], lay.numel[b]});
      o = matvec_labelled(x, W, Lenv, Renv, cs);
    }
    // grow the layout by the output's new blocks
    std::vector<int> where(o.blocks.size());
    bool grow = false;
    for (size_t q = 0; q < o.blocks.size(); ++q) {
      where[q] = o.blocks[q].numel ? lay.find(o.blocks[q].idx) : 0;
      grow |= where[q] < 0;
    }
    if (grow) {
      Layout nl;
      nl.rank = lay.rank;
      nl.idx = lay.idx;
      nl.numel = lay.numel;
      for (size_t q = 0; q < o.blocks.size(); ++q)
        if (where[q] < 0) nl.add(o.blocks[q].idx, o.blocks[q].numel);
      nl.finish();
      BufPtr oldV = V;
      const int64_t old_n = lay.n;
      const int64_t oldv = oldV->dptr() - ab;
      alloc_vectors(nl.n);
      std::vector<kern::Copy2D> cp;
      std::vector<kern::ChunkTask> tasks;
      std::vector<char> np(nl.idx.size(), 0);
      for (size_t b = 0; b < nl.idx.size(); ++b) {
        const int k = lay.find(nl.idx[b]);
        if (k >= 0) np[b] = present[k];
        for (int r = 0; r <= j; ++r) {
          add_chunks(tasks, int(cp.size()), nl.numel[b]);
          cp.push_back(flat_copy(k >= 0 ? oldv + int64_t(r) * old_n + lay.off[k] : -1,
                                 V->dptr() - ab + int64_t(r) * nl.n + nl.off[b], nl.numel[b]));
        }
      }
      launch_copies(ctx, ab, cp, tasks, "lanczos_regrow");
      lay = nl;
      present.swap(np);
      for (size_t q = 0; q < o.blocks.size(); ++q) where[q] = o.blocks[q].numel ? lay.find(o.blocks[q].idx) : 0;
    }
    // w = H v_j in the layout (absent blocks zero)
    {
      std::vector<kern::Copy2D> cp;
      std::vector<kern::ChunkTask> tasks;
      std::vector<char> hit(lay.idx.size(), 0);
      for (size_t q = 0; q < o.blocks.size(); ++q) {
        if (!o.blocks[q].numel) continue;
        const int k = where[q];
        hit[k] = 1;
        add_chunks(tasks, int(cp.size()), lay.numel[k]);
        cp.push_back(flat_copy(o.ptr(o.blocks[q]) - ab, woff() + lay.off[k], lay.numel[k]));
      }
      for (size_t k = 0; k < lay.idx.size(); ++k)
        if (!hit[k]) {
          add_chunks(tasks, int(cp.size()), lay.numel[k]);
          cp.push_back(flat_copy(-1, woff() + lay.off[k], lay.numel[k]));
        }
      launch_copies(ctx, ab, cp, tasks, "lanczos_scatter");
    }
    {
      // two classical Gram-Schmidt passes against v_0..v_j, then |w|^2 and
      // the tridiagonal eigensolve (alpha_j = <v_j|H v_j> with its correction)
      KernelTimer kt(*ctx, "lanczos_vec");
      const int mj = j + 1;
      kern::krylov_dot(ctx->stream(), ab, voff(0), lay.n, mj, woff(), lay.n, poff(), c1);
      kern::krylov_axpy(ctx->stream(), ab, voff(0), lay.n, mj, c1, -1.0, woff(), woff(), lay.n, -1);
      kern::krylov_dot(ctx->stream(), ab, voff(0), lay.n, mj, woff(), lay.n, poff(), c2);
      kern::krylov_axpy(ctx->stream(), ab, voff(0), lay.n, mj, c2, -1.0, woff(), woff(), lay.n, -1);
      kern::krylov_dot(ctx->stream(), ab, woff(), lay.n, 1, woff(), lay.n, poff(), nrm2);
    }
    {
      KernelTimer kt(*ctx, "lanczos_tridiag");
      kern::lanczos_tridiag(ctx->stream(), ab, t_off, K, j, c1, c2, nrm2, out, d_flag);
    }
    double host_all[3 + kern::kMaxKrylov + 1];
    ctx->read_small(host_all, ab + out, size_t(3 + K + 1) * sizeof(double));
    int32_t hflag = 0;
    std::memcpy(&hflag, host_all + 3 + K, sizeof(int32_t));
    if (hflag) throw LogicError("nan found in output tensor");
    std::copy(host_all, host_all + 3, host_out);
    const double lam = host_out[0], resid = host_out[1], beta = host_out[2];
    m = j + 1;
    const bool done = resid <= tol * std::fabs(lam) || !(beta > 1e-15 * std::fabs(lam)) || m == K;
    if (done) break;
    for (size_t q = 0; q < o.blocks.size(); ++q)
      if (o.blocks[q].numel) present[where[q]] = 1;
    KernelTimer kt(*ctx, "lanczos_vec");
    kern::krylov_axpy(ctx->stream(), ab, voff(0), lay.n, 0, nrm2, 0.0, woff(), voff(j + 1), lay.n, nrm2);
  }
  res.iterations = m;
  res.graph_replays = graph.replays;
  res.energy = host_out[0];
  res.residual = host_out[1];
  // theta' = sum_i y_i v_i, restricted to the blocks the Krylov vectors have
  {
    KernelTimer kt(*ctx, "lanczos_vec");
    kern::krylov_axpy(ctx->stream(), ab, voff(0), lay.n, m, out + 3, 1.0, -1, woff(), lay.n, -1);
  }
  BTensor t2(ctx, theta.shape);
  std::vector<BIdx> keep;
  for (size_t k = 0; k < lay.idx.size(); ++k)
    if (present[k]) keep.push_back(lay.idx[k]);
  t2.allocate(keep);
  {
    std::vector<kern::Copy2D> cp;
    std::vector<kern::ChunkTask> tasks;
    for (auto& b : t2.blocks) {
      const int k = lay.find(b.idx);
      add_chunks(tasks, int(cp.size()), b.numel);
      cp.push_back(flat_copy(woff() + lay.off[k], t2.ptr(b) - ab, b.numel));
    }
    launch_copies(ctx, ab, cp, tasks, "lanczos_scatter");
  }
  res.theta = std::move(t2);
  Shape rank0;
  rank0.g = theta.shape.g;
  rank0.sel = q_neutral();
  res.E = BTensor(ctx, rank0);
  res.E.storage = sc;
  res.E.blocks.push_back({BIdx{}, out - (sc->dptr() - ab), 1});
  return res;
}

}  // namespace qtt
