// Block linear algebra: sector grouping of btensor_linalg.cpp on the host,
// batched Jacobi SVD / eigh and the global truncation on device.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <functional>
#include <limits>
#include <numeric>

#include "btensor.hpp"
#include "kernels.hpp"

namespace qtt {

namespace {

double* arena_base(Context* c) { return reinterpret_cast<double*>(c->arena().base()); }

void add_chunks(std::vector<kern::ChunkTask>& tasks, int desc, int64_t numel) {
  const int64_t nch = (numel + kern::kChunk - 1) / kern::kChunk;
  for (int64_t c = 0; c < nch; ++c) tasks.push_back({desc, int32_t(c)});
}

struct Slice {
  int sec;
  int64_t start, size;
};
struct SectorPlan {
  std::vector<int> blocks;  // indices into the rank-2 tensor's block list
  std::vector<Slice> rows;  // encounter order
  std::vector<Slice> cols;  // sorted by section index, offsets in encounter order
  int64_t m = 0, n = 0;
};

// reorder_by_cvals + compact_dense + compact_dense_single
// (sources/btensor_linalg.cpp:30-255) for a rank-2 tensor.
std::vector<SectorPlan> compact_dense_plan(const BTensor& t) {
  const Shape& s = t.shape;
  std::vector<int> ord(t.blocks.size());
  std::iota(ord.begin(), ord.end(), 0);
  auto rc = [&](int k) { return s.cvals[0][t.blocks[k].idx.i[0]]; };
  auto cc = [&](int k) { return s.cvals[1][t.blocks[k].idx.i[1]]; };
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
    const Charge ra = rc(a), rb = rc(b);
    if (ra < rb) return true;
    return ra == rb && cc(a) < cc(b);
  });
  std::vector<SectorPlan> out;
  size_t st = 0;
  while (st < ord.size()) {
    size_t en = st + 1;
    while (en < ord.size() && rc(ord[en]) == rc(ord[st]) && cc(ord[en]) == cc(ord[st])) ++en;
    SectorPlan p;
    int cur_row = t.blocks[ord[st]].idx.i[0];
    p.rows.push_back({cur_row, 0, s.secs[0][cur_row]});
    p.m = s.secs[0][cur_row];
    for (size_t q = st; q < en; ++q) {
      const BIdx& bi = t.blocks[ord[q]].idx;
      p.blocks.push_back(ord[q]);
      const int c = bi.i[1];
      auto it = std::lower_bound(p.cols.begin(), p.cols.end(), c, [](const Slice& x, int v) { return x.sec < v; });
      if (it == p.cols.end() || it->sec != c) {
        p.cols.insert(it, Slice{c, p.n, s.secs[1][c]});
        p.n += s.secs[1][c];
      }
      if (bi.i[0] != cur_row) {
        cur_row = bi.i[0];
        p.rows.push_back({cur_row, p.m, s.secs[0][cur_row]});
        p.m += s.secs[0][cur_row];
      }
    }
    out.push_back(std::move(p));
    st = en;
  }
  return out;
}

struct Batched {
  std::vector<kern::SvdSector> secs;
  BufPtr work, outbuf;
  std::vector<int64_t> u_off, v_off, s_off;  // element offsets into outbuf
  int sweeps = 0;
  int unconverged = 0;  // sectors still rotating when a sweep cap was reached
  std::vector<char> wide;  // sectors whose rotations ran as multi-work-group rounds (k_svd_round)
  // defer_complete: the orthonormal completion of null columns, one task per
  // sector (scratch offsets unset), left to the caller
  std::vector<kern::CompleteTask> complete;
};

// A grouped GEMM whose descriptors are on the device (one staging copy), so
// it can be launched several times.
struct GemmLaunch {
  const kern::GemmTile* tiles = nullptr;
  const kern::GemmOut* outs = nullptr;
  const kern::GemmPair* pairs = nullptr;
  int ntiles = 0;
};

GemmLaunch prepare_gemm(Context* ctx, const std::vector<kern::GemmOut>& outs,
                        const std::vector<kern::GemmPair>& pairs) {
  std::vector<kern::GemmTile> tiles;
  for (size_t o = 0; o < outs.size(); ++o) {
    const int tm = (outs[o].M + kern::kGemmTile - 1) / kern::kGemmTile;
    const int tn = (outs[o].N + kern::kGemmTile - 1) / kern::kGemmTile;
    for (int i = 0; i < tm; ++i)
      for (int j = 0; j < tn; ++j) tiles.push_back({int32_t(o), int16_t(i), int16_t(j)});
  }
  GemmLaunch g;
  if (tiles.empty()) return g;
  std::tie(g.tiles, g.outs, g.pairs) = ctx->upload_pack(tiles, outs, pairs);
  g.ntiles = int(tiles.size());
  return g;
}

void run_gemm(Context* ctx, double* ab, const GemmLaunch& g, const char* name, double alpha = 1.0,
              double beta = 0.0) {
  if (g.ntiles == 0) return;
  KernelTimer kt(*ctx, name);
  kern::gemm_f64(ctx->stream(), ab, g.tiles, g.ntiles, g.outs, g.pairs, alpha, beta);
}

void launch_gemm(Context* ctx, double* ab, const std::vector<kern::GemmOut>& outs,
                 const std::vector<kern::GemmPair>& pairs, const char* name, double alpha = 1.0,
                 double beta = 0.0) {
  run_gemm(ctx, ab, prepare_gemm(ctx, outs, pairs), name, alpha, beta);
}

int env_int(const char* name, int dflt) {
  const char* v = std::getenv(name);
  return v ? std::atoi(v) : dflt;
}

// Cyclic block Jacobi sweeps over the sectors listed in `which` until a sweep
// rotates nothing (or max_sweeps).  first_rot[i] receives the number of
// rotating work-groups of the first sweep of sector which[i].  Returns the
// number of sweeps run; *capped (if given) is raised by the number of
// sectors still rotating when max_sweeps ran out.
//
// The sweeps are pipelined one deep while they rotate much: sweep s+1 is
// enqueued before the host waits for the rotation counts of sweep s, so the
// device does not idle on that read.  Sweep s+1 is built from the counts up
// to sweep s-1; its work-groups of a sector that sweep s left unrotated
// return at once (skip pointer), so every sector stops after the same sweeps
// as with a read between sweeps.  Once a sweep rotates under a quarter of its
// pair-rounds the end is near and the next sweep is enqueued only after the
// read (a speculative sweep after the last would be launches of no-ops).
//
// The sweep that ends the loop rotates nothing: it only confirms that every
// column pair passes the round test.  Once a sweep rotates clearly fewer
// pair-rounds than the sweeps before it (the tail of the quadratic
// convergence), the sweep after it is followed by that test on all pairs at
// once instead of a speculative sweep: G = Y^T Y for the active sectors (one
// grouped GEMM) and the round's relative test on every pair of G
// (k_gram_check).  A sector whose pairs all pass has converged: the sweep
// after would have rotated nothing.
int jacobi_sweeps(Context* ctx, double* ab, const std::vector<kern::SvdSector>& secs, const std::vector<int>& which,
                  double tol, int inner, int max_sweeps, std::vector<int>* first_rot, int* capped) {
  const int ns = int(secs.size());
  if (which.empty() || max_sweeps <= 0) return 0;
  BufPtr cnt = ctx->alloc(size_t(4 * ns) * sizeof(int32_t));  // counts of two sweeps, check flags
  int32_t* d_cnt[2] = {reinterpret_cast<int32_t*>(cnt->cptr()), reinterpret_cast<int32_t*>(cnt->cptr()) + ns};
  int32_t* d_chk = reinterpret_cast<int32_t*>(cnt->cptr()) + 2 * ns;
  ctx->reserve_read(size_t(ns));
  std::vector<char> active(ns, 0);
  for (int i : which) active[i] = 1;
  if (first_rot) first_rot->assign(which.size(), 0);
  // Rounds after the first of a sweep rotate only the pairs across their two
  // blocks (the first round holds every block once and takes all its pairs,
  // so a sweep still meets every column pair).  QTT_SVD_CROSS_INNER: 1 (the
  // default) from the first sweep on, 2 from the second, 0 never.
  const int cross = env_int("QTT_SVD_CROSS_INNER", 1);
  // enqueues sweep `sweep` over the active sectors and the read of its counts
  std::vector<int64_t> slot_rounds(2, 0);  // pair-rounds enqueued per slot
  // The work items of every sector are uploaded once: a sector that has
  // converged keeps its items, whose work-groups return at once (skip
  // pointer), and the rounds launched cover the active sectors' widest.
  std::vector<kern::RoundWork> work;
  int max_ell = 0;
  for (int i : which) {
    const auto& sc = secs[i];
    max_ell = std::max(max_ell, sc.ell);
    for (int pi = 0; pi < sc.nb / 2; ++pi) work.push_back({sc.y_off, sc.w_off, sc.ell, sc.nb, sc.flag_idx, pi});
  }
  // longest rows first: when a round has more work-groups than the device
  // holds at once, the long ones start in the first wave of work-groups
  std::stable_sort(work.begin(), work.end(),
                   [](const kern::RoundWork& a, const kern::RoundWork& b) { return a.ell > b.ell; });
  const kern::RoundWork* d_work = work.empty() ? nullptr : ctx->upload_vec(work);
  auto enqueue = [&](int sweep) {
    int max_nb = 0;
    int64_t nactive = 0;
    for (int i = 0; i < ns; ++i) {
      if (!active[i]) continue;
      max_nb = std::max(max_nb, secs[i].nb);
      nactive += secs[i].nb / 2;
    }
    if (nactive == 0) return false;
    const int slot = sweep & 1;
    slot_rounds[size_t(slot)] = nactive * (max_nb - 1);
    const int* skip = sweep > 0 ? d_cnt[slot ^ 1] : nullptr;
    kern::fill_i32(ctx->stream(), d_cnt[slot], ns, 0);
    for (int r = 0; r < max_nb - 1; ++r) {
      KernelTimer kt(*ctx, "svd_round");
      kern::svd_round(ctx->stream(), ab, d_work, int(work.size()), r, d_cnt[slot], skip, tol, inner,
                      max_ell, cross == 1 || (cross == 2 && sweep > 0));
    }
    QTT_HIP(hipMemcpyAsync(ctx->read_slot(slot), d_cnt[slot], size_t(ns) * sizeof(int32_t), hipMemcpyDeviceToHost,
                           ctx->stream()));
    QTT_HIP(hipEventRecord(ctx->read_event(slot), ctx->stream()));
    return true;
  };
  // the all-pairs test after the sweep just enqueued (flags: failing pairs,
  // relative and absolute, of sector i at d_chk[2 i], d_chk[2 i + 1])
  BufPtr gram;
  auto enqueue_check = [&] {
    std::vector<kern::GemmOut> outs;
    std::vector<kern::GemmPair> prs;
    std::vector<kern::SmallJacobi> cd;
    int64_t elems = 0;
    int kc = 0;
    for (int i = 0; i < ns; ++i)
      if (active[i]) elems += int64_t(secs[i].k) * secs[i].k;
    if (elems == 0) return false;
    gram = ctx->alloc(size_t(elems) * sizeof(double));
    int64_t go = gram->dptr() - ab;
    for (int i = 0; i < ns; ++i) {
      if (!active[i]) continue;
      const auto& sc = secs[i];
      kern::GemmPair p{};  // G = Y^T Y over the k columns (Y: ell x k, column-major)
      p.a_off = sc.y_off;
      p.a_rs = sc.ell;
      p.a_cs = 1;
      p.b_off = sc.y_off;
      p.b_rs = 1;
      p.b_cs = sc.ell;
      p.K = sc.ell;
      outs.push_back({go, sc.k, sc.k, int32_t(prs.size()), int32_t(prs.size() + 1), sc.k, 0});
      prs.push_back(p);
      cd.push_back({go, -1, sc.k, sc.k, sc.k, sc.k, 2 * i, 2 * i + 1});
      kc = std::max(kc, sc.k);
      go += int64_t(sc.k) * sc.k;
    }
    launch_gemm(ctx, ab, outs, prs, "svd_sweep_check_gram");
    KernelTimer kt(*ctx, "svd_sweep_check");
    kern::gram_check(ctx->stream(), ab, ctx->upload_vec(cd), int(cd.size()), d_chk, tol, kc);
    return true;
  };
  const bool use_check = env_int("QTT_SVD_SWEEP_CHECK", 1) != 0;
  if (!enqueue(0)) return 0;
  int queued = 1, sweeps = 0;
  bool speculate = true, check_next = false;
  double frac_max = 0.0;
  std::vector<int32_t> h_chk(size_t(2 * ns)), h_zero;
  for (int sweep = 0; sweep < queued; ++sweep) {
    bool checked = false;
    if (queued == sweep + 1) {
      if (speculate && queued < max_sweeps && enqueue(queued))
        ++queued;
      else if (check_next && enqueue_check())
        checked = true;
    }
    ctx->wait_event(ctx->read_event(sweep & 1));
    const int32_t* h_cnt = ctx->read_slot(sweep & 1);
    sweeps = sweep + 1;
    if (sweep == 0 && first_rot)
      for (size_t q = 0; q < which.size(); ++q) (*first_rot)[q] = h_cnt[which[q]];
    bool any = false;
    int64_t rotated = 0;
    for (int i = 0; i < ns; ++i) {
      rotated += h_cnt[i];
      if (active[i] && h_cnt[i] == 0) active[i] = 0;
      any |= bool(active[i]);
    }
    if (any && checked) {
      ctx->read_small(h_chk.data(), d_chk, size_t(2 * ns) * sizeof(int32_t));
      h_zero.assign(h_cnt, h_cnt + ns);
      bool passed = false;
      any = false;
      for (int i = 0; i < ns; ++i) {
        if (active[i] && h_chk[size_t(2 * i)] + h_chk[size_t(2 * i + 1)] == 0) {
          active[i] = 0;
          h_zero[size_t(i)] = 0;  // the next sweep's work-groups of this sector return at once
          passed = true;
        }
        any |= bool(active[i]);
      }
      if (any && passed) ctx->stage_to(d_cnt[sweep & 1], h_zero.data(), size_t(ns) * sizeof(int32_t));
    }
    if (!any) break;
    const double frac = double(rotated) / double(std::max<int64_t>(slot_rounds[size_t(sweep & 1)], 1));
    frac_max = std::max(frac_max, frac);
    speculate = 4 * rotated >= slot_rounds[size_t(sweep & 1)] && (!use_check || frac >= 0.75 * frac_max);
    check_next = use_check && !speculate;
    if (queued == sweep + 1 && queued < max_sweeps && enqueue(queued)) ++queued;
  }
  if (capped && sweeps >= max_sweeps)
    for (int i = 0; i < ns; ++i) *capped += active[i];
  return sweeps;
}

// Gathers the sector matrices of a rank-2 tensor and runs the batched Jacobi
// SVD (eigh_mode: shifted symmetric eigen-decomposition) on device.
//
// Sectors much longer than wide (ell >= 2 kp) are first preconditioned by
// Gram iterations: G = Y^T Y (kp x kp, MFMA GEMM), its eigenvectors W_g by the
// same block Jacobi on G, Y <- Y W_g and W <- W W_g, repeated until the
// Jacobi on G rotates nothing in its first sweep, i.e. the columns of Y are
// orthogonal to the tolerance.  Every step is an exact orthogonal transform
// of Y, so the singular values still come from the column norms of Y.
Batched run_batched(const BTensor& t, const std::vector<SectorPlan>& plans, bool eigh_mode,
                    bool defer_complete = false) {
  Context* ctx = t.ctx;
  double* ab = arena_base(ctx);
  Batched B;
  const int ns = int(plans.size());
  B.secs.resize(ns);
  int64_t work_elems = 0, out_elems = 0;
  std::vector<int64_t> yo(ns), wo(ns), wsz(ns);
  B.u_off.resize(ns);
  B.v_off.resize(ns);
  B.s_off.resize(ns);
  for (int i = 0; i < ns; ++i) {
    const auto& p = plans[i];
    if (eigh_mode && p.m != p.n) throw InvalidArgument("eigh requires square sectors");
    auto& s = B.secs[i];
    s.m = int32_t(p.m);
    s.n = int32_t(p.n);
    s.transposed = p.m < p.n;
    s.ell = int32_t(std::max(p.m, p.n));
    s.k = int32_t(std::min(p.m, p.n));
    const int nblk = std::max(1, (s.k + kern::kSvdBlock - 1) / kern::kSvdBlock);
    s.nb = 2 * ((nblk + 1) / 2);
    const int64_t kp = int64_t(s.nb) * kern::kSvdBlock;
    s.flag_idx = i;
    s.eig_shift = eigh_mode;
    yo[i] = work_elems;
    work_elems += int64_t(s.ell) * kp;
    wo[i] = work_elems;
    wsz[i] = kp * kp + 2 * kp + s.ell + 1;  // W + finalize scratch + eigh shift
    work_elems += wsz[i];
    B.u_off[i] = out_elems;
    out_elems += int64_t(s.m) * s.k;
    B.v_off[i] = out_elems;
    out_elems += eigh_mode ? 0 : int64_t(s.n) * s.k;
    B.s_off[i] = out_elems;
    out_elems += s.k;
  }
  B.work = ctx->alloc(size_t(std::max<int64_t>(work_elems, 1)) * sizeof(double));
  B.outbuf = ctx->alloc(size_t(std::max<int64_t>(out_elems, 1)) * sizeof(double));
  const int64_t wb = B.work->dptr() - ab, ob = B.outbuf->dptr() - ab;
  for (int i = 0; i < ns; ++i) {
    auto& s = B.secs[i];
    const int64_t kp = int64_t(s.nb) * kern::kSvdBlock;
    s.y_off = wb + yo[i];
    s.w_off = wb + wo[i];
    s.shift_off = s.w_off + kp * kp + 2 * kp + s.ell;
    s.u_off = ob + B.u_off[i];
    s.v_off = ob + B.v_off[i];
    s.s_off = ob + B.s_off[i];
  }
  // gather: zero the Y matrices, then copy every block into its sector slot
  std::vector<kern::Copy2D> cp;
  std::vector<kern::ChunkTask> tasks;
  for (int i = 0; i < ns; ++i) {
    const auto& s = B.secs[i];
    const int64_t kp = int64_t(s.nb) * kern::kSvdBlock;
    kern::Copy2D z{};
    z.src_off = -1;
    z.dst_off = s.y_off;
    z.rows = 1;
    z.cols = int32_t(int64_t(s.ell) * kp);
    z.dst_rs = 0;
    z.dst_cs = 1;
    add_chunks(tasks, int(cp.size()), int64_t(z.cols));
    cp.push_back(z);
  }
  if (!tasks.empty()) {
    KernelTimer kt(*ctx, "svd_zero");
    auto [dc, dt] = ctx->upload_pack(cp, tasks);
    kern::copy2d(ctx->stream(), ab, dc, dt, int(tasks.size()));
  }
  cp.clear();
  tasks.clear();
  for (int i = 0; i < ns; ++i) {
    const auto& p = plans[i];
    const auto& s = B.secs[i];
    for (int bk : p.blocks) {
      const Block& b = t.blocks[bk];
      if (!b.numel) continue;
      const auto rit = std::find_if(p.rows.begin(), p.rows.end(), [&](const Slice& x) { return x.sec == b.idx.i[0]; });
      const auto cit = std::find_if(p.cols.begin(), p.cols.end(), [&](const Slice& x) { return x.sec == b.idx.i[1]; });
      kern::Copy2D c{};
      c.src_off = t.ptr(b) - ab;
      c.rows = int32_t(rit->size);
      c.cols = int32_t(cit->size);
      c.src_rs = cit->size;
      c.src_cs = 1;
      c.alpha = 1.0;
      if (!s.transposed) {  // Y = A (col-major, ld = m)
        c.dst_off = s.y_off + rit->start + cit->start * int64_t(s.ell);
        c.dst_rs = 1;
        c.dst_cs = s.ell;
      } else {  // Y = A^T (col-major, ld = n)
        c.dst_off = s.y_off + cit->start + rit->start * int64_t(s.ell);
        c.dst_rs = s.ell;
        c.dst_cs = 1;
      }
      add_chunks(tasks, int(cp.size()), b.numel);
      cp.push_back(c);
    }
  }
  if (!tasks.empty()) {
    KernelTimer kt(*ctx, "svd_gather");
    auto [dc, dt] = ctx->upload_pack(cp, tasks);
    kern::copy2d(ctx->stream(), ab, dc, dt, int(tasks.size()));
  }
  if (ns == 0) return B;
  {
    auto* d_secs = ctx->upload_vec(B.secs);
    KernelTimer kt(*ctx, "svd_init");
    kern::svd_init_w(ctx->stream(), ab, d_secs, ns);
    if (eigh_mode) kern::svd_shift(ctx->stream(), ab, d_secs, ns);
  }
  int max_ell = 0;
  for (auto& s : B.secs) max_ell = std::max(max_ell, s.ell);
  const double tol = 8.0 * std::sqrt(double(std::max(max_ell, 16))) * DBL_EPSILON;
  const int inner = env_int("QTT_SVD_INNER_SWEEPS", 1);
  const int max_sweeps = env_int("QTT_SVD_MAX_SWEEPS", 30);
  const size_t lds_limit = kern::jacobi_small_lds_limit();
  const size_t lds_margin = 256;
  auto lds_bytes = [](int64_t ell, int64_t k) { return size_t(ell * (k + (k & 1))) * sizeof(double); };
  const bool use_small = env_int("QTT_SVD_SMALL", 1) != 0;
  std::vector<char> done(ns, 0), big_only(ns, 0);
  B.wide.assign(ns, 0);
  std::vector<BufPtr> keep;  // Gram and ping-pong buffers live until the finalize is enqueued

  // ---- Gram iterations ----------------------------------------------------------
  // Every SVD sector goes through G = Y^T Y, eigenvectors W_g of G, Y <- Y W_g,
  // W <- W W_g until G is diagonal to the tolerance (relative, or absolute
  // with respect to the largest singular value; the second case ends with
  // two Gram-Schmidt passes on the short columns).  Sectors with k <=
  // kGramEigMaxK solve for W_g in one work-group (LDS); wider ones run the
  // pivoted Cholesky by panels, block Jacobi rounds on L = R^T and a
  // normalisation pass, all with G in device memory.
  const size_t pchol_limit = kern::pchol_lds_limit();
  std::vector<int> wide;
  std::vector<char> large(ns, 0);
  const int eig_max_k = std::min(kern::kGramEigMaxK, env_int("QTT_SVD_EIG_MAX_K", kern::kGramEigMaxK));
  for (int i = 0; i < ns; ++i) {
    const auto& s = B.secs[i];
    if (eigh_mode || !use_small || s.k < 1) continue;
    if (s.k <= eig_max_k && kern::gram_eig_lds_bytes(s.k) + lds_margin <= lds_limit) {
      wide.push_back(i);
    } else if (size_t(s.k) * (4 + kern::kPcholExtra) * sizeof(double) <= pchol_limit &&
               env_int("QTT_SVD_LARGE_GRAM", 1)) {
      wide.push_back(i);
      large[i] = 1;
      B.wide[i] = 1;
    }
  }
  if (!wide.empty()) {
    const int nw = int(wide.size());
    int kmax_large = 0;
    for (int i : wide)
      if (large[i]) kmax_large = std::max(kmax_large, B.secs[i].k);
    const int pnb =
        kmax_large
            ? int(std::min<size_t>(32, pchol_limit / (size_t(kmax_large) * sizeof(double)) - kern::kPcholExtra))
            : 0;
    int64_t elems = 0;
    std::vector<int64_t> go(nw), gwo(nw), y2o(nw), w2o(nw), lo(nw, -1), lpo(nw, -1), sco(nw, -1), sto(nw, -1);
    size_t lds_need = 0;
    for (int q = 0; q < nw; ++q) {
      const auto& s = B.secs[wide[q]];
      const int64_t kp = int64_t(s.nb) * kern::kSvdBlock;
      const int64_t k = s.k;
      go[q] = elems;
      elems += k * k;
      gwo[q] = elems;
      elems += k * k;
      y2o[q] = elems;
      elems += int64_t(s.ell) * kp;
      w2o[q] = elems;
      elems += wsz[wide[q]];
      if (large[wide[q]]) {
        lo[q] = elems;  // L, k x kp with zero padding columns
        elems += k * kp;
        lpo[q] = elems;
        elems += k * pnb;
        sco[q] = elems;
        elems += 3 * k;
        sto[q] = elems;  // ints
        elems += (k + 2) / 2 + 1;
      } else {
        lds_need = std::max(lds_need, kern::gram_eig_lds_bytes(s.k));
      }
    }
    BufPtr gbuf = ctx->alloc(size_t(elems) * sizeof(double));
    keep.push_back(gbuf);
    const int64_t gb = gbuf->dptr() - ab;
    {  // ping-pong partners: Y2 zero (padding columns stay zero), W2 identity
      std::vector<kern::Copy2D> z;
      std::vector<kern::ChunkTask> zt;
      std::vector<kern::SvdSector> wi;
      for (int q = 0; q < nw; ++q) {
        const auto& s = B.secs[wide[q]];
        const int64_t kp = int64_t(s.nb) * kern::kSvdBlock;
        kern::Copy2D c{};
        c.src_off = -1;
        c.dst_off = gb + y2o[q];
        c.rows = 1;
        c.cols = int32_t(int64_t(s.ell) * kp);
        c.dst_cs = 1;
        add_chunks(zt, int(z.size()), int64_t(c.cols));
        z.push_back(c);
        kern::SvdSector w = s;
        w.w_off = gb + w2o[q];
        wi.push_back(w);
      }
      KernelTimer kt(*ctx, "svd_init");
      auto [dz, dzt] = ctx->upload_pack(z, zt);
      kern::copy2d(ctx->stream(), ab, dz, dzt, int(zt.size()));
      kern::svd_init_w(ctx->stream(), ab, ctx->upload_vec(wi), nw);
    }
    std::vector<int64_t> alt_y(nw), alt_w(nw);
    for (int q = 0; q < nw; ++q) {
      alt_y[q] = gb + y2o[q];
      alt_w[q] = gb + w2o[q];
    }
    BufPtr fbuf = ctx->alloc(size_t(2 * nw) * sizeof(int32_t));
    keep.push_back(fbuf);
    int32_t* d_flags = reinterpret_cast<int32_t*>(fbuf->cptr());
    std::vector<int32_t> h_flags(size_t(2 * nw), 0);
    std::vector<int> act(nw), corr;
    std::iota(act.begin(), act.end(), 0);
    auto gram_of = [&](const std::vector<int>& qs) {  // G = Y^T Y over the k real columns
      // Long sectors split K over several tiles (partials summed in a fixed
      // order afterwards) so that few small tiles do not serialise ell rows.
      std::vector<kern::GemmOut> outs;
      std::vector<kern::GemmPair> prs;
      std::vector<int> parts(qs.size(), 1);
      int64_t pelems = 0;
      for (size_t j = 0; j < qs.size(); ++j) {
        const auto& s = B.secs[wide[qs[j]]];
        parts[j] = int(std::min<int64_t>(16, std::max<int64_t>(1, s.ell / env_int("QTT_GRAM_ROWS_PER_PART", 256))));
        if (parts[j] > 1) pelems += int64_t(parts[j]) * s.k * s.k;
      }
      BufPtr pbuf = pelems ? ctx->alloc(size_t(pelems) * sizeof(double)) : nullptr;
      const int64_t pb = pbuf ? pbuf->dptr() - ab : 0;
      std::vector<kern::SumParts> sums;
      std::vector<kern::ChunkTask> st;
      int64_t pcur = pb;
      for (size_t j = 0; j < qs.size(); ++j) {
        const int q = qs[j];
        const auto& s = B.secs[wide[q]];
        const int np = parts[j];
        const int64_t kc = (int64_t(s.ell) + np - 1) / np;
        for (int p = 0; p < np; ++p) {
          const int64_t k0 = p * kc, k1 = std::min<int64_t>(s.ell, k0 + kc);
          kern::GemmPair gp{};
          gp.a_off = s.y_off + k0;
          gp.a_rs = s.ell;
          gp.a_cs = 1;
          gp.b_off = s.y_off + k0;
          gp.b_rs = 1;
          gp.b_cs = s.ell;
          gp.K = int32_t(std::max<int64_t>(0, k1 - k0));
          const int64_t c_off = np > 1 ? pcur + int64_t(p) * s.k * s.k : gb + go[q];
          outs.push_back({c_off, s.k, s.k, int32_t(prs.size()), int32_t(prs.size() + 1), s.k, 0});
          prs.push_back(gp);
        }
        if (np > 1) {
          add_chunks(st, int(sums.size()), int64_t(s.k) * s.k);
          sums.push_back({gb + go[q], pcur, int64_t(s.k) * s.k, np, 0});
          pcur += int64_t(np) * s.k * s.k;
        }
      }
      launch_gemm(ctx, ab, outs, prs, "svd_gram");
      if (!st.empty()) {
        KernelTimer kt(*ctx, "svd_gram_sum");
        auto [ds, dst] = ctx->upload_pack(sums, st);
        kern::sum_parts(ctx->stream(), ab, ds, dst, int(st.size()));
      }
    };
    auto read_flags = [&] { ctx->read_small(h_flags.data(), d_flags, size_t(2 * nw) * sizeof(int32_t)); };
    auto retire = [&](std::vector<int>& qs) {  // drop converged sectors after a flag read
      std::vector<int> next;
      for (int q : qs) {
        if (h_flags[q] == 0) {
          done[wide[q]] = 1;
          if (h_flags[nw + q] > 0) corr.push_back(q);  // converged in the absolute sense only
        } else {
          next.push_back(q);
        }
      }
      qs.swap(next);
    };
    auto rotate = [&](const std::vector<int>& qs) {  // Y2^T = W_g^T Y^T, W2^T = W_g^T W^T
      std::vector<kern::GemmOut> outs;
      std::vector<kern::GemmPair> prs;
      for (int q : qs) {
        const auto& s = B.secs[wide[q]];
        const int kp = s.nb * kern::kSvdBlock;
        kern::GemmPair p1{};
        p1.a_off = gb + gwo[q];
        p1.a_rs = s.k;
        p1.a_cs = 1;
        p1.b_off = s.y_off;
        p1.b_rs = s.ell;
        p1.b_cs = 1;
        p1.K = s.k;
        outs.push_back({alt_y[q], s.k, s.ell, int32_t(prs.size()), int32_t(prs.size() + 1), s.ell, 0});
        prs.push_back(p1);
        kern::GemmPair p2 = p1;
        p2.b_off = s.w_off;
        p2.b_rs = kp;
        outs.push_back({alt_w[q], s.k, s.k, int32_t(prs.size()), int32_t(prs.size() + 1), kp, 0});
        prs.push_back(p2);
      }
      launch_gemm(ctx, ab, outs, prs, "svd_rotate");
      for (int q : qs) {
        auto& s = B.secs[wide[q]];
        std::swap(s.y_off, alt_y[q]);
        std::swap(s.w_off, alt_w[q]);
      }
    };
    auto solve_large = [&](const std::vector<int>& qs) {  // W_g of wide sectors, G in device memory
      if (qs.empty()) return;
      {
        std::vector<kern::Copy2D> z;
        std::vector<kern::ChunkTask> zt;
        for (int q : qs) {
          const auto& s = B.secs[wide[q]];
          const int64_t kp = int64_t(s.nb) * kern::kSvdBlock;
          kern::Copy2D c{};
          c.src_off = -1;
          c.dst_off = gb + lo[q];
          c.rows = 1;
          c.cols = int32_t(int64_t(s.k) * kp);
          c.dst_cs = 1;
          add_chunks(zt, int(z.size()), int64_t(c.cols));
          z.push_back(c);
        }
        KernelTimer kt(*ctx, "svd_init");
        auto [dz, dzt] = ctx->upload_pack(z, zt);
        kern::copy2d(ctx->stream(), ab, dz, dzt, int(zt.size()));
      }
      int kact = 0;
      for (int q : qs) kact = std::max(kact, B.secs[wide[q]].k);
      // Descriptors of every panel step, uploaded once: a sector whose
      // factorisation has ended sees a zero panel and a null update.
      std::vector<kern::CholPanel> pd;
      std::vector<kern::GemmOut> outs;
      std::vector<kern::GemmPair> prs;
      for (int q : qs) {
        const auto& s = B.secs[wide[q]];
        pd.push_back({gb + go[q], gb + lo[q], gb + lpo[q], gb + sto[q], s.k, 0});
        kern::GemmPair gp{};  // G -= Lp Lp^T
        gp.a_off = gb + lpo[q];
        gp.a_rs = 1;
        gp.a_cs = s.k;
        gp.b_off = gb + lpo[q];
        gp.b_rs = s.k;
        gp.b_cs = 1;
        gp.K = pnb;
        outs.push_back({gb + go[q], s.k, s.k, int32_t(prs.size()), int32_t(prs.size() + 1), s.k, 0});
        prs.push_back(gp);
      }
      const auto* d_pd = ctx->upload_vec(pd);
      kern::pchol_init(ctx->stream(), ab, d_pd, int(pd.size()));
      const GemmLaunch upd = prepare_gemm(ctx, outs, prs);
      for (int j0 = 0; j0 < kact; j0 += pnb) {
        {
          KernelTimer kt(*ctx, "svd_pchol");
          kern::pchol_panel(ctx->stream(), ab, d_pd, int(pd.size()), j0, pnb, kact);
        }
        if (j0 + pnb < kact) run_gemm(ctx, ab, upd, "svd_pchol_update", -1.0, 1.0);
      }
      std::vector<kern::SvdSector> gsec;
      std::vector<int> all;
      std::vector<kern::GramVec> vd;
      for (int q : qs) {
        const auto& s = B.secs[wide[q]];
        kern::SvdSector g = s;
        g.y_off = gb + lo[q];
        g.w_off = -1;  // the eigenvectors come from the normalised columns
        g.ell = s.k;
        g.flag_idx = int32_t(gsec.size());
        all.push_back(int(gsec.size()));
        gsec.push_back(g);
        vd.push_back({gb + lo[q], gb + gwo[q], gb + sco[q], s.k, s.k, s.k, 0});
      }
      jacobi_sweeps(ctx, ab, gsec, all, tol, inner, max_sweeps, nullptr, &B.unconverged);
      int kv = 0;
      for (auto& v : vd) kv = std::max(kv, v.k);
      KernelTimer kt(*ctx, "svd_vectors");
      kern::gram_vectors(ctx->stream(), ab, ctx->upload_vec(vd), int(vd.size()), kv);
    };
    const int max_it = env_int("QTT_SVD_GRAM_ITERS", 6);
    for (int it = 0; it < max_it && !act.empty(); ++it) {
      gram_of(act);
      std::vector<int> act_s, act_l;
      for (int q : act) (large[wide[q]] ? act_l : act_s).push_back(q);
      // small sectors: test and solve in one work-group each.  Beside large
      // sectors they run on the side stream, concurrently with the large
      // sectors' Cholesky and rounds (which leave most CUs free); the main
      // stream joins before the rotation.  Profiling keeps one stream so that
      // the event timers stay per kernel.
      const bool overlap = !act_s.empty() && !act_l.empty() && !ctx->profiling && ctx->overlap;
      if (!act_s.empty()) {
        std::vector<kern::SmallJacobi> jd;
        for (int q : act_s) {
          const auto& s = B.secs[wide[q]];
          jd.push_back({gb + go[q], gb + gwo[q], s.k, s.k, s.k, s.k, q, nw + q});
        }
        // widest first, so that the longest solves start in the first wave of
        // work-groups when there are more sectors than the device holds
        std::stable_sort(jd.begin(), jd.end(),
                         [](const kern::SmallJacobi& a, const kern::SmallJacobi& b) { return a.k > b.k; });
        static const char* const kEigNames[] = {"svd_eig", "svd_eig_it1", "svd_eig_it2", "svd_eig_it3+"};
        const auto* d_jd = ctx->upload_vec(jd);
        if (overlap) ctx->fork();
        KernelTimer kt(*ctx, kEigNames[std::min(it, 3)]);
        kern::gram_eig(overlap ? ctx->side_stream() : ctx->stream(), ab, d_jd, int(jd.size()), d_flags, tol,
                       max_sweeps, env_int("QTT_SVD_PRECOND", 1), lds_need);
      }
      // large sectors: test first (the solve is a long chain of launches); the
      // flag read also covers the small solver launched above
      int cd_launched = 0;
      if (!act_l.empty()) {
        cd_launched = 1;
        std::vector<kern::SmallJacobi> cd;
        int kc = 0;
        for (int q : act_l) {
          const auto& s = B.secs[wide[q]];
          cd.push_back({gb + go[q], -1, s.k, s.k, s.k, s.k, q, nw + q});
          kc = std::max(kc, s.k);
        }
        {
          KernelTimer kt(*ctx, "svd_check");
          kern::gram_check(ctx->stream(), ab, ctx->upload_vec(cd), int(cd.size()), d_flags, tol, kc);
        }
        read_flags();  // the small sectors' flags may not be written yet when overlapped
        retire(act_l);
        solve_large(act_l);
      }
      if (overlap) ctx->join();
      if (it >= 1) {  // the small solver's test of this iteration's G
        if (!act_s.empty() && ((act_l.empty() && cd_launched == 0) || overlap)) read_flags();
        retire(act_s);
      }
      act = act_s;
      act.insert(act.end(), act_l.begin(), act_l.end());
      if (act.empty()) break;
      rotate(act);
      B.sweeps += 1;
    }
    for (int q : act) big_only[wide[q]] = 1;  // keeps its accumulated W: multi-WG path only
    // Sectors converged with residual couplings below the absolute floor:
    // two first-order Gram-Schmidt passes Y <- Y T (longer columns first)
    // make the short columns orthogonal; W is unchanged, and the dropped
    // triangular part is below tol times the largest singular value.
    for (int pass = 0; pass < 2 && !corr.empty(); ++pass) {
      if (pass > 0) gram_of(corr);  // the first pass reuses the last Gram matrix of the loop
      std::vector<kern::SmallJacobi> gd;
      int kg = 0;
      for (int q : corr) {
        const auto& s = B.secs[wide[q]];
        gd.push_back({gb + go[q], gb + gwo[q], s.k, s.k, s.k, s.k, -1, -1});
        kg = std::max(kg, s.k);
      }
      {
        KernelTimer kt(*ctx, "svd_gs");
        kern::gs_step(ctx->stream(), ab, ctx->upload_vec(gd), int(gd.size()), kg);
      }
      std::vector<kern::GemmOut> outs;
      std::vector<kern::GemmPair> prs;
      for (int q : corr) {
        const auto& s = B.secs[wide[q]];
        kern::GemmPair p1{};
        p1.a_off = gb + gwo[q];
        p1.a_rs = s.k;
        p1.a_cs = 1;
        p1.b_off = s.y_off;
        p1.b_rs = s.ell;
        p1.b_cs = 1;
        p1.K = s.k;
        outs.push_back({alt_y[q], s.k, s.ell, int32_t(prs.size()), int32_t(prs.size() + 1), s.ell, 0});
        prs.push_back(p1);
      }
      launch_gemm(ctx, ab, outs, prs, "svd_rotate");
      for (int q : corr) std::swap(B.secs[wide[q]].y_off, alt_y[q]);
    }
  }
  // ---- Jacobi on everything not yet orthogonal --------------------------------
  std::vector<int> rest_small, rest_big;
  size_t small_need = 0;
  for (int i = 0; i < ns; ++i) {
    if (done[i]) continue;
    const auto& s = B.secs[i];
    if (use_small && !big_only[i] && lds_bytes(s.ell, s.k) + lds_margin <= lds_limit) {
      rest_small.push_back(i);
      small_need = std::max(small_need, lds_bytes(s.ell, s.k));
    } else {
      rest_big.push_back(i);
    }
  }
  if (!rest_small.empty()) {
    std::vector<kern::SmallJacobi> jd;
    for (int i : rest_small) {
      const auto& s = B.secs[i];
      jd.push_back({s.y_off, s.w_off, s.ell, s.k, s.ell, s.nb * kern::kSvdBlock, -1, 0});
    }
    KernelTimer kt(*ctx, "svd_small");
    kern::jacobi_small(ctx->stream(), ab, ctx->upload_vec(jd), int(jd.size()), nullptr, tol, max_sweeps,
                       small_need);
  }
  for (int i : rest_big) B.wide[i] = 1;
  B.sweeps += jacobi_sweeps(ctx, ab, B.secs, rest_big, tol, inner, max_sweeps, nullptr, &B.unconverged);
  // ---- finalize --------------------------------------------------------------
  std::vector<int4> wt;
  for (int i = 0; i < ns; ++i) {
    const auto& s = B.secs[i];
    if (!eigh_mode)
      for (int r = 0; r < s.ell; r += 64)
        for (int c = 0; c < s.k; c += 64) wt.push_back(make_int4(i, 0, r, c));
    for (int r = 0; r < s.k; r += 64)
      for (int c = 0; c < s.k; c += 64) wt.push_back(make_int4(i, 1, r, c));
  }
  {
    KernelTimer kt(*ctx, "svd_finalize");
    kern::svd_finalize(ctx->stream(), ab, ctx->upload_vec(B.secs), B.secs.data(), ns, ctx->upload_vec(wt),
                       int(wt.size()));
  }
  if (!eigh_mode) {
    for (const auto& s : B.secs) {
      const int64_t kp = int64_t(s.nb) * kern::kSvdBlock;
      kern::CompleteTask c{};
      c.z_off = s.transposed ? s.v_off : s.u_off;
      c.s_off = s.s_off;
      c.ell = s.ell;
      c.k = s.k;
      c.limit = s.k;
      // scratch in the sector's work area: W has been written out by now
      c.p_off = s.w_off;
      c.v_off = s.w_off + kp * kp + 2 * kp;
      B.complete.push_back(c);
    }
    if (!defer_complete) {
      KernelTimer kt(*ctx, "svd_complete");
      kern::svd_complete(ctx->stream(), ab, ctx->upload_vec(B.complete), int(B.complete.size()));
      B.complete.clear();
    }
  }
  // the work buffers are released after the finalize kernels in stream order
  return B;
}

}  // namespace

// svd(const btensor&, some, uv) on the rank-2 reshaped tensor + reshape_as
// back (sources/btensor_linalg.cpp:390-534).
// deferred (may be null): receives the completion tasks of the null columns
// instead of running them (one per sector = section of d).
static SvdResult svd_impl(const BTensor& t, size_t split, SvdStats* stats,
                          std::vector<kern::CompleteTask>* deferred) {
  Context* ctx = t.ctx;
  if (split > size_t(t.rank())) throw InvalidArgument("split larger than the rank");
  BTensor rt = reshape(t, {int64_t(split)});
  const Shape& rs = rt.shape;
  const Groups& g = rs.g;
  auto plans = compact_dense_plan(rt);
  Batched B = run_batched(rt, plans, false, deferred != nullptr);
  if (deferred) *deferred = B.complete;
  const int ns = int(plans.size());
  // d (neutral), U (rows, k) with the row charges and rule = rule of t,
  // V (cols, k) with inverse charges and neutral rule
  std::vector<int64_t> dsz(ns);
  std::vector<Charge> rd_c(ns), ld_c(ns);
  for (int i = 0; i < ns; ++i) {
    dsz[i] = B.secs[i].k;
    rd_c[i] = rs.cvals[1][plans[i].cols[0].sec];
    ld_c[i] = q_inv(g, rd_c[i]);
  }
  Shape ds;
  ds.g = g;
  ds.secs = {dsz};
  ds.cvals = {std::vector<Charge>(ns, q_neutral())};
  ds.sel = q_neutral();
  Shape us;
  us.g = g;
  us.secs = {rs.secs[0], dsz};
  us.cvals = {rs.cvals[0], rd_c};
  us.sel = rs.sel;
  Shape vs;
  vs.g = g;
  vs.secs = {rs.secs[1], dsz};
  vs.cvals = {rs.cvals[1], ld_c};
  vs.sel = q_neutral();
  vs = vs.inverse_cvals();
  BTensor d(ctx, ds), U(ctx, us), V(ctx, vs);
  d.storage = U.storage = V.storage = B.outbuf;
  for (int i = 0; i < ns; ++i) {
    const int64_t k = B.secs[i].k;
    BIdx di{};
    di.i[0] = i;
    d.blocks.push_back({di, B.s_off[i], k});
    for (auto& r : plans[i].rows) {
      BIdx ui{};
      ui.i[0] = r.sec;
      ui.i[1] = i;
      U.blocks.push_back({ui, B.u_off[i] + r.start * k, r.size * k});
    }
    for (auto& c : plans[i].cols) {
      BIdx vi{};
      vi.i[0] = c.sec;
      vi.i[1] = i;
      V.blocks.push_back({vi, B.v_off[i] + c.start * k, c.size * k});
    }
  }
  auto srt = [](BTensor& x) {
    const int r = x.rank();
    std::sort(x.blocks.begin(), x.blocks.end(),
              [r](const Block& a, const Block& b) { return bidx_less(a.idx, b.idx, r); });
  };
  srt(U);
  srt(V);
  for (size_t i = 0; i < B.secs.size(); ++i) {
    const auto& s = B.secs[i];
    const double m = std::max(s.m, s.n), n = std::min(s.m, s.n);
    const double f = 14.0 * m * n * n + 8.0 * n * n * n;
    ctx->svd_model_flops += f;
    if (i < B.wide.size() && B.wide[i]) ctx->svd_model_flops_wide += f;
  }
  if (stats) {
    for (auto& s : B.secs) {
      const double m = std::max(s.m, s.n), n = std::min(s.m, s.n);
      stats->flops += 14.0 * m * n * n + 8.0 * n * n * n;
      stats->max_k = std::max(stats->max_k, s.k);
    }
    stats->sectors += ns;
    stats->sweeps = std::max(stats->sweeps, B.sweeps);
    stats->unconverged += B.unconverged;
  }
  ctx->svd_calls += 1;
  ctx->svd_max_sweeps = std::max<int64_t>(ctx->svd_max_sweeps, B.sweeps);
  ctx->svd_unconverged += B.unconverged;
  // reshape_as back to the original legs
  std::vector<int64_t> ushape(t.rank(), -1), vshape(t.rank(), -1);
  for (int i = 0; i < t.rank(); ++i) (i < int(split) ? vshape : ushape)[i] = 0;
  Shape u_left = t.shape.shape_from(ushape);
  Shape u_right = U.shape.shape_from({0, -1});
  Shape v_left = t.shape.shape_from(vshape).inverse_cvals();
  Shape v_right = V.shape.shape_from({0, -1});
  SvdResult out;
  out.U = reshape_as(U, shape_product({&u_left, &u_right}));
  out.V = reshape_as(V, shape_product({&v_left, &v_right}));
  out.d = d;
  return out;
}

SvdResult svd(const BTensor& t, size_t split, SvdStats* stats) { return svd_impl(t, split, stats, nullptr); }

// eigh(const btensor&, split) (sources/btensor_linalg.cpp:294-389): shifted
// symmetric Jacobi, eigenvalues ascending per sector.
EighResult eigh(const BTensor& t, size_t split) {
  Context* ctx = t.ctx;
  BTensor rt = reshape(t, {int64_t(split)});
  const Shape& rs = rt.shape;
  const Groups& g = rs.g;
  auto plans = compact_dense_plan(rt);
  Batched B = run_batched(rt, plans, true);
  ctx->svd_calls += 1;
  ctx->svd_max_sweeps = std::max<int64_t>(ctx->svd_max_sweeps, B.sweeps);
  ctx->svd_unconverged += B.unconverged;
  const int ns = int(plans.size());
  std::vector<int64_t> dsz(ns);
  std::vector<Charge> rd_c(ns);
  for (int i = 0; i < ns; ++i) {
    dsz[i] = B.secs[i].k;
    rd_c[i] = rs.cvals[1][plans[i].cols[0].sec];
  }
  Shape ds;
  ds.g = g;
  ds.secs = {dsz};
  ds.cvals = {std::vector<Charge>(ns, q_neutral())};
  ds.sel = q_neutral();
  Shape us;
  us.g = g;
  us.secs = {rs.secs[0], dsz};
  us.cvals = {rs.cvals[0], rd_c};
  us.sel = q_add(g, rs.shape_from({-1, 0}).sel, q_inv(g, rs.cvals[1].at(0)));
  BTensor d(ctx, ds), U(ctx, us);
  d.storage = U.storage = B.outbuf;
  for (int i = 0; i < ns; ++i) {
    const int64_t k = B.secs[i].k;
    BIdx di{};
    di.i[0] = i;
    d.blocks.push_back({di, B.s_off[i], k});
    for (auto& r : plans[i].rows) {
      BIdx ui{};
      ui.i[0] = r.sec;
      ui.i[1] = i;
      if (!us.allowed(ui))
        throw InvalidArgument("block index " + fmt_idx(ui, 2) + " not allowed by selection rule");
      U.blocks.push_back({ui, B.u_off[i] + r.start * k, r.size * k});
    }
  }
  const int r = 2;
  std::sort(U.blocks.begin(), U.blocks.end(), [r](const Block& a, const Block& b) { return bidx_less(a.idx, b.idx, r); });
  std::vector<int64_t> ushape(t.rank(), -1);
  for (int i = int(split); i < t.rank(); ++i) ushape[i] = 0;
  Shape u_left = t.shape.shape_from(ushape);
  Shape u_right = U.shape.shape_from({0, -1});
  EighResult out;
  out.U = reshape_as(U, shape_product({&u_left, &u_right}));
  out.d = d;
  return out;
}

namespace {

// remove_unit_blocks lambda of truncate_impl (sources/btensor_linalg.cpp:690-709),
// step for step: it advances past one matching block per iteration, so a
// matching block that directly follows another matching block survives.
void remove_unit_blocks(std::vector<Block>& lst, int rank, int kd) {
  size_t src = 0, dest = 0;
  const size_t n = lst.size();
  while (dest != n) {
    const int dest_last = lst[dest].idx.i[rank - 1];
    dest += (dest_last == kd);
    if (dest != src && dest != n) std::swap(lst[dest], lst[src]);
    dest += (dest != n);
    ++src;
  }
  lst.resize(src);
}

}  // namespace

namespace {

// truncate_impl (sources/btensor_linalg.cpp:657-755) with the cut computed on
// device, for d and any number of unitaries whose last leg carries d's
// sections (U and V of an svd, the eigenvectors of an eigh).  The unitaries
// keep their block order, sliced blocks are compacted into fresh storage;
// *units are replaced and the truncated d is returned.
// before_copy (may be empty) runs once the kept counts are known, before any
// block is copied, with the kept count of every section of d (0 for sections
// without a block).
BTensor truncate_units(BTensor&& d_in, const std::vector<BTensor*>& units, size_t max_size, size_t min_size,
                       double tol, double pw,
                       const std::function<void(const std::vector<int32_t>&)>& before_copy = {}) {
  BTensor d = std::move(d_in);
  Context* ctx = d.ctx;
  double* ab = arena_base(ctx);
  const int nd = int(d.blocks.size());
  if (nd == 0) return d;
  std::vector<int64_t> s_offs(nd), s_pos(nd);
  std::vector<int32_t> ks(nd);
  int64_t total = 0;
  for (int i = 0; i < nd; ++i) {
    s_offs[i] = d.ptr(d.blocks[i]) - ab;
    ks[i] = int32_t(d.blocks[i].numel);
    s_pos[i] = total;
    total += ks[i];
  }
  if (total == 0) return d;
  int64_t n2 = 1;
  while (n2 < total) n2 <<= 1;
  BufPtr scratch = ctx->alloc(size_t(n2 + 1) * sizeof(double) + size_t(nd) * sizeof(int32_t) + 256);
  int32_t* d_kept = reinterpret_cast<int32_t*>(scratch->cptr() + size_t(n2 + 1) * sizeof(double));
  double* d_thr = scratch->dptr() + n2;
  const int64_t clamp = std::numeric_limits<int64_t>::max();
  {
    KernelTimer kt(*ctx, "truncate_select");
    auto [up_s_offs, up_ks, up_pos] = ctx->upload_pack(s_offs, ks, s_pos);
    kern::truncate_select(ctx->stream(), ab, up_s_offs, up_ks, up_pos, nd,
                          scratch->dptr() - ab, total, tol, pw, int64_t(std::min<size_t>(min_size, size_t(clamp))),
                          int64_t(std::min<size_t>(max_size, size_t(clamp))), d_kept, d_thr);
  }
  std::vector<int32_t> kept(nd);
  ctx->read_small(kept.data(), d_kept, size_t(nd) * sizeof(int32_t));
  if (before_copy) {
    std::vector<int32_t> kept_sec(d.shape.secs[0].size(), 0);
    for (int i = 0; i < nd; ++i) kept_sec[d.blocks[i].idx.i[0]] = kept[i];
    before_copy(kept_sec);
  }

  const size_t nu = units.size();
  std::vector<std::vector<Block>> ub(nu);
  std::vector<Shape> ush(nu);
  for (size_t u = 0; u < nu; ++u) {
    ub[u] = units[u]->blocks;
    ush[u] = units[u]->shape;
  }
  std::vector<Block> db = d.blocks;
  Shape dsh = d.shape;
  std::vector<char> d_erased(nd, 0);
  for (int i = nd - 1; i >= 0; --i) {
    const int kd = db[i].idx.i[0];
    if (kept[i] == 0) {
      for (size_t u = 0; u < nu; ++u) remove_unit_blocks(ub[u], units[u]->rank(), kd);
      d_erased[i] = 1;
    } else {
      dsh.secs[0][kd] = kept[i];
      for (size_t u = 0; u < nu; ++u) ush[u].secs[units[u]->rank() - 1][kd] = kept[i];
    }
  }
  auto rebuild = [&](const BTensor& src, const std::vector<Block>& keep, const Shape& nsh, int r) {
    BTensor out(ctx, nsh);
    std::vector<BIdx> idx;
    for (auto& b : keep) idx.push_back(b.idx);
    // keep the (possibly quirk-reordered) order of the reference's list
    out.blocks.clear();
    int64_t off = 0;
    for (auto& i : idx) {
      const int64_t n = nsh.block_numel(i);
      out.blocks.push_back({i, off, n});
      off += n;
    }
    out.storage = ctx->alloc(size_t(std::max<int64_t>(off, 1)) * sizeof(double));
    std::vector<kern::Copy2D> cp;
    std::vector<kern::ChunkTask> tasks;
    for (size_t q = 0; q < keep.size(); ++q) {
      const Block& sb = keep[q];
      const Block& nbk = out.blocks[q];
      if (!nbk.numel) continue;
      int64_t dims[kMaxRank];
      src.shape.block_dims(sb.idx, dims);
      const int64_t oldk = dims[r - 1];
      const int64_t rows = oldk ? sb.numel / oldk : 0;
      const int64_t newk = nsh.secs[r - 1][sb.idx.i[r - 1]];
      kern::Copy2D c{};
      c.src_off = src.ptr(sb) - ab;
      c.dst_off = out.ptr(nbk) - ab;
      c.rows = int32_t(rows);
      c.cols = int32_t(newk);
      c.src_rs = oldk;
      c.src_cs = 1;
      c.dst_rs = newk;
      c.dst_cs = 1;
      c.alpha = 1.0;
      add_chunks(tasks, int(cp.size()), rows * newk);
      cp.push_back(c);
    }
    if (!tasks.empty()) {
      KernelTimer kt(*ctx, "truncate_copy");
      auto [dc, dt] = ctx->upload_pack(cp, tasks);
      kern::copy2d(ctx->stream(), ab, dc, dt, int(tasks.size()));
    }
    return out;
  };
  // a surviving unitary block of an erased sector keeps its stale size: the
  // section size of an erased sector is never changed
  std::vector<Block> dkeep;
  for (int i = 0; i < nd; ++i)
    if (!d_erased[i]) dkeep.push_back(db[i]);
  for (size_t u = 0; u < nu; ++u) *units[u] = rebuild(*units[u], ub[u], ush[u], units[u]->rank());
  return rebuild(d, dkeep, dsh, 1);
}

}  // namespace

SvdResult truncate(SvdResult&& in, size_t max_size, size_t min_size, double tol, double pw) {
  SvdResult out = std::move(in);
  out.d = truncate_units(std::move(out.d), {&out.U, &out.V}, max_size, min_size, tol, pw);
  return out;
}

// truncate(std::tuple<btensor, btensor>&& e_S, ...) (sources/btensor_linalg.cpp:786-800):
// the eigenvalues take the place of the singular values and the eigenvectors
// are truncated with them.  (The reference passes the eigenvalue tensor a
// second time where the eigenvectors belong, btensor_linalg.cpp:797, so its
// second output is a moved-from tensor; the eigenvectors are what the
// declaration and the truncating eigh above it mean.)  Eigenvalues come in
// ascending order within a sector, and a sector keeps the values before the
// first one not above the cut (lower_bound_impl2, btensor_linalg.cpp:545-555):
// a sector whose smallest eigenvalue is at or below the cut is erased whole.
EighResult truncate(EighResult&& in, size_t max_size, size_t min_size, double tol, double pw) {
  EighResult out = std::move(in);
  out.d = truncate_units(std::move(out.d), {&out.U}, max_size, min_size, tol, pw);
  return out;
}

EighResult eigh_truncated(const BTensor& t, size_t split, double tol, size_t min_size, size_t max_size, double pw) {
  return truncate(eigh(t, split), max_size, min_size, tol, pw);
}

// svd(const btensor&, BOOL some, BOOL compute_uv) (sources/btensor_linalg.cpp:390-502)
// of a rank-2 tensor.  With some or compute_uv false the reference sizes the
// unitaries' last sections by max(m, n) (or torch::svd returns m x m and
// n x n zero matrices) while d keeps min(m, n) values per sector; the
// structure is consistent only for square sectors, where the result equals
// the economy form.  Rectangular sectors are rejected in those modes.
SvdResult svd_matrix(const BTensor& t, bool some, bool compute_uv) {
  if (t.rank() != 2) throw InvalidArgument("svd(tensor, some, compute_uv) takes a rank-2 tensor");
  if (!some || !compute_uv) {
    for (const auto& p : compact_dense_plan(t))
      if (p.m != p.n)
        throw InvalidArgument("svd with some = false or compute_uv = false: sector of " + std::to_string(p.m) + " x " +
                              std::to_string(p.n) + " is not square (the reference's d, U and V sizes disagree there)");
  }
  SvdResult r = svd(t, 1);
  if (!compute_uv) {  // zero-filled U and V of the same structure
    Context* ctx = t.ctx;
    double* ab = arena_base(ctx);
    std::vector<kern::Copy2D> z;
    std::vector<kern::ChunkTask> zt;
    for (BTensor* u : {&r.U, &r.V}) {
      for (const auto& b : u->blocks) {
        if (!b.numel) continue;
        kern::Copy2D c{};
        c.src_off = -1;
        c.dst_off = u->ptr(b) - ab;
        c.rows = 1;
        c.cols = int32_t(b.numel);
        c.dst_cs = 1;
        add_chunks(zt, int(z.size()), b.numel);
        z.push_back(c);
      }
    }
    if (!zt.empty()) {
      KernelTimer kt(*ctx, "svd_zero_uv");
      auto [dz, dzt] = ctx->upload_pack(z, zt);
      kern::copy2d(ctx->stream(), ab, dz, dzt, int(zt.size()));
    }
  }
  return r;
}

// The truncation drops every column whose singular value is numerically zero
// unless min_size keeps it, so the orthonormal completion of such columns runs
// after the cut and only for the columns that are kept.
SvdResult svd_truncated(const BTensor& t, size_t split, double tol, size_t min_size, size_t max_size, double pw,
                        SvdStats* stats) {
  std::vector<kern::CompleteTask> pend;
  SvdResult out = svd_impl(t, split, stats, &pend);
  Context* ctx = t.ctx;
  auto complete_kept = [&](const std::vector<int32_t>& kept_sec) {
    std::vector<kern::CompleteTask> run;
    int64_t elems = 0;
    for (size_t i = 0; i < pend.size() && i < kept_sec.size(); ++i) {
      if (kept_sec[i] <= 0 || pend[i].k <= 0) continue;
      kern::CompleteTask c = pend[i];
      c.limit = kept_sec[i];
      c.v_off = elems;  // relative to the scratch buffer until it exists
      elems += c.ell;
      c.p_off = elems;
      elems += c.k;
      run.push_back(c);
    }
    if (run.empty()) return;
    BufPtr scratch = ctx->alloc(size_t(elems) * sizeof(double));
    const int64_t so = scratch->dptr() - arena_base(ctx);
    for (auto& c : run) {
      c.v_off += so;
      c.p_off += so;
    }
    KernelTimer kt(*ctx, "svd_complete");
    kern::svd_complete(ctx->stream(), arena_base(ctx), ctx->upload_vec(run), int(run.size()));
  };
  out.d = truncate_units(std::move(out.d), {&out.U, &out.V}, max_size, min_size, tol, pw, complete_kept);
  return out;
}

}  // namespace qtt
