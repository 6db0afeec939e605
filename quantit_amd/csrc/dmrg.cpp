// Two-site DMRG with one-step Lanczos (sources/dmrg.cpp), every tensor
// operation on device; the host keeps the sector bookkeeping and reads back
// only the kept-dimension counts of each truncation and one energy per sweep.
#include "dmrg.hpp"

#include <chrono>
#include <cmath>

#include "kernels.hpp"

namespace qtt {

// edge_shape_prep (dmrg.cpp:334-342).  The arguments of its final shape_from
// are evaluated right to left (g++), so the rank-0 part is taken before
// inverse_cvals_.
Shape edge_shape_prep(const Shape& t, int dim) {
  std::vector<int64_t> vec(t.rank(), 0);
  vec[dim] = -1;
  Shape S = t.shape_from(vec);
  S.sel = q_neutral();
  Shape s0 = S.shape_from({0});
  Shape Si = S.inverse_cvals();
  return shape_product({&Si, &s0});
}

namespace {
// Labels the contraction GEMMs of a call site in the profiling timers.
struct GemmLabel {
  Context* ctx;
  const char* prev;
  GemmLabel(Context* c, const char* name) : ctx(c), prev(c->gemm_label) { ctx->gemm_label = name; }
  ~GemmLabel() { ctx->gemm_label = prev; }
};
}  // namespace

BTensor compute_left_env(const BTensor& W, const BTensor& Y, const BTensor& L) {
  GemmLabel gl(L.ctx, "gemm_env");
  BTensor out = tensordot(L, Y, {0}, {0});
  out = tensordot(out, W, {0, 2}, {0, 3});
  return tensordot(out, conj(Y), {0, 2}, {0, 1});
}

BTensor compute_right_env(const BTensor& W, const BTensor& Y, const BTensor& R) {
  GemmLabel gl(R.ctx, "gemm_env");
  BTensor out = tensordot(R, Y, {0}, {2});
  out = tensordot(out, W, {0, 3}, {2, 3});
  return tensordot(out, conj(Y), {3, 0}, {1, 2});
}

BTensor hamil2site_times_state(const BTensor& theta, const BTensor& W12, const BTensor& L, const BTensor& R) {
  BTensor out = tensordot(L, theta, {0}, {0});
  out = tensordot(out, W12, {0, 2, 3}, {0, 4, 5});
  return tensordot(out, R, {1, 4}, {0, 1});
}

namespace {

BTensor& env_at(DmrgState& st, int i) { return st.env.at(size_t(i + 1)); }

int64_t bond_of(const BTensor& a) {
  int64_t n = 0;
  for (auto x : a.shape.secs[2]) n += x;
  return n;
}

// rank-0 tensor holding 0 when a Lanczos scalar has no block (no matching
// block pair): the reference's arithmetic then runs on an empty tensor.
BTensor zero_scalar(Context* ctx, const Groups& g) {
  Shape s;
  s.g = g;
  s.sel = q_neutral();
  return fill_like(ctx, s, 0.0);
}

// eye_like of a rank-2 shape: identity on the diagonal blocks.
BTensor eye_like(Context* ctx, const Shape& s) {
  std::vector<BIdx> idx;
  std::vector<double> host;
  for (int i = 0; i < s.nsec(0); ++i)
    for (int j = 0; j < s.nsec(1); ++j) {
      BIdx b{};
      b.i[0] = i;
      b.i[1] = j;
      if (!s.allowed(b)) continue;
      const int64_t r = s.secs[0][i], c = s.secs[1][j];
      if (i != j || r != c) continue;
      idx.push_back(b);
      for (int64_t x = 0; x < r; ++x)
        for (int64_t y = 0; y < c; ++y) host.push_back(x == y ? 1.0 : 0.0);
    }
  BTensor out(ctx, s);
  out.upload(idx, host.data());
  return out;
}

}  // namespace

void dmrg_prepare(DmrgState& st) {
  const int L = int(st.mpo.size());
  if (L < 2 || int(st.mps.size()) != L) throw InvalidArgument("MPS and MPO must have the same length >= 2");
  if (st.oc < 0 || st.oc >= L) throw InvalidArgument("orthogonality center outside the MPS");
  Context* ctx = st.mps[0].ctx;
  st.env.assign(size_t(L + 2), BTensor());
  Shape Ls = edge_shape_prep(st.mps.front().shape, 0);
  Shape LH = edge_shape_prep(st.mpo.front().shape, 0);
  Shape Lsi = Ls.inverse_cvals();
  env_at(st, -1) = fill_like(ctx, shape_product({&Ls, &LH, &Lsi}), 1.0);
  Shape Rs = edge_shape_prep(st.mps.back().shape, 2);
  Shape RH = edge_shape_prep(st.mpo.back().shape, 2);
  Shape Rsi = Rs.inverse_cvals();
  env_at(st, L) = fill_like(ctx, shape_product({&Rs, &RH, &Rsi}), 1.0);
  // generate_env_impl (dmrg.cpp:370-409): left environments below the
  // orthogonality centre, right environments above it
  for (int i = 0; i < st.oc; ++i) env_at(st, i) = compute_left_env(st.mpo[i], st.mps[i], env_at(st, i - 1));
  for (int i = L - 1; i > st.oc; --i) env_at(st, i) = compute_right_env(st.mpo[i], st.mps[i], env_at(st, i + 1));
  st.w12.clear();
  for (int i = 0; i + 1 < L; ++i)
    st.w12.push_back(permute(tensordot(st.mpo[i], st.mpo[i + 1], {2}, {0}), {0, 1, 3, 4, 2, 5}));
}

BTensor matvec_labelled(const BTensor& x, const BTensor& W, const BTensor& Lenv, const BTensor& Renv,
                        ContractStats* cs) {
  Context* ctx = x.ctx;
  BTensor out;
  {
    GemmLabel gl(ctx, "gemm_mv_L");
    out = tensordot(Lenv, x, {0}, {0}, cs);
  }
  {
    GemmLabel gl(ctx, "gemm_mv_W");
    out = tensordot(out, W, {0, 2, 3}, {0, 4, 5}, cs);
  }
  GemmLabel gl(ctx, "gemm_mv_R");
  return tensordot(out, Renv, {1, 4}, {0, 1}, cs);
}

LanczosStep one_step_lanczos(const BTensor& theta, const BTensor& W, const BTensor& Lenv, const BTensor& Renv,
                             ContractStats* cs) {
  Context* ctx = theta.ctx;
  const Groups& g = theta.shape.g;
  LanczosStep r;
  BTensor psi = matvec_labelled(theta, W, Lenv, Renv, cs);
  r.a0 = full_dot(psi, theta);
  if (r.a0.blocks.empty()) r.a0 = zero_scalar(ctx, g);
  else psi = add(psi, mul(theta, r.a0), -1.0);
  r.b = full_dot(psi, psi, true);
  if (r.b.blocks.empty()) r.b = zero_scalar(ctx, g);
  r.Hpsi = div_scalar_tensor(psi, r.b, 1e-15);
  r.a1 = full_dot(matvec_labelled(r.Hpsi, W, Lenv, Renv, cs), r.Hpsi);
  if (r.a1.blocks.empty()) r.a1 = zero_scalar(ctx, g);
  return r;
}

Eig2x2 eig2x2_dev(const BTensor& a0, const BTensor& a1, const BTensor& b) {
  Context* ctx = a0.ctx;
  Shape s0;
  s0.g = a0.shape.g;
  s0.sel = q_neutral();
  BufPtr trio = ctx->alloc(4 * sizeof(double));
  double* ab = reinterpret_cast<double*>(ctx->arena().base());
  Eig2x2 r{BTensor(ctx, s0), BTensor(ctx, s0), BTensor(ctx, s0), nullptr};
  for (BTensor* x : {&r.E, &r.o, &r.n}) {
    x->storage = trio;
    x->blocks.push_back({BIdx{}, 0, 1});
  }
  r.o.blocks[0].off = 1;
  r.n.blocks[0].off = 2;
  r.flag = ctx->alloc(sizeof(int));
  kern::fill_i32(ctx->stream(), reinterpret_cast<int32_t*>(r.flag->cptr()), 1, 0);
  KernelTimer kt(*ctx, "eig2x2");
  kern::eig2x2(ctx->stream(), ab, a0.ptr(a0.blocks[0]) - ab, a1.ptr(a1.blocks[0]) - ab, b.ptr(b.blocks[0]) - ab,
               trio->dptr() - ab, reinterpret_cast<int*>(r.flag->cptr()));
  return r;
}

// The post-SVD part of the update (dmrg.cpp:188-201): d normalised, absorbed
// into V (forward) or U (backward), environment of the site left behind.
static void absorb(DmrgState& st, SvdResult& r, int oc) {
  normalize_d(r.d);
  if (st.step == 1) {
    st.mps[oc] = r.U;
    st.mps[oc + 1] = permute(conj(mul(r.V, r.d)), {2, 0, 1});
    env_at(st, oc) = compute_left_env(st.mpo[oc], st.mps[oc], env_at(st, oc - 1));
  } else {
    st.mps[oc] = mul(r.U, r.d);
    st.mps[oc + 1] = permute(conj(r.V), {2, 0, 1});
    env_at(st, oc + 1) = compute_right_env(st.mpo[oc + 1], st.mps[oc + 1], env_at(st, oc + 2));
  }
}

BTensor dmrg_update(DmrgState& st, const DmrgOptions& opt, DmrgLog* log) {
  const int oc = st.oc;
  Context* ctx = st.mps[0].ctx;
  ContractStats cs;
  SvdStats ss;
  BTensor& Lenv = env_at(st, oc - 1);
  BTensor& Renv = env_at(st, oc + 2);
  const BTensor& W = st.w12[oc];
  BTensor theta = tensordot(st.mps[oc], st.mps[oc + 1], {2}, {0}, &cs);
  if (ctx->lanczos_krylov >= 2) {
    LanczosResult lr = lanczos_ground(theta, W, Lenv, Renv, ctx->lanczos_krylov, ctx->lanczos_tol, &cs);
    SvdResult r = svd_truncated(lr.theta, 2, opt.cutoff, opt.minimum_bond, opt.maximum_bond, 2.0, &ss);
    absorb(st, r, oc);
    if (log) {
      log->contract_flops += cs.flops;
      log->svd_flops += ss.flops;
      log->updates += 1;
      log->lanczos_vectors += lr.iterations;
    }
    ctx->lanczos_updates += 1;
    ctx->lanczos_vectors += lr.iterations;
    ctx->graph_replays += lr.graph_replays;
    return lr.E;
  }
  // one_step_lanczos (dmrg.cpp:584-608), eig2x2Mat (dmrg.cpp:543-572) on device
  LanczosStep ls = one_step_lanczos(theta, W, Lenv, Renv, &cs);
  Eig2x2 eg = eig2x2_dev(ls.a0, ls.a1, ls.b);
  const BTensor& psi = ls.Hpsi;
  const BTensor& E = eg.E;
  const BTensor& o = eg.o;
  const BTensor& n = eg.n;
  // the flag lands in pinned memory without a wait; the truncation's read of
  // its kept counts synchronizes the stream after it
  int32_t* h_flag = ctx->pinned_flag();
  *h_flag = 0;
  QTT_HIP(hipMemcpyAsync(h_flag, eg.flag->cptr(), sizeof(int), hipMemcpyDeviceToHost, ctx->stream()));
  BTensor theta2 = axpby_dev(theta, o, psi, n);
  SvdResult r = svd_truncated(theta2, 2, opt.cutoff, opt.minimum_bond, opt.maximum_bond, 2.0, &ss);
  ctx->sync();  // normally already done by the truncation; the flag copy is complete
  if (*h_flag) throw LogicError("nan found in output tensor");
  absorb(st, r, oc);
  if (log) {
    log->contract_flops += cs.flops;
    log->svd_flops += ss.flops;
    log->updates += 1;
  }
  return E;
}

double dmrg_sweep(DmrgState& st, const DmrgOptions& opt, DmrgLog* log) {
  const int L = int(st.mpo.size());
  const int nstep = int(st.w12.size()) - 1 + (st.w12.size() == 1);
  BTensor E;
  for (int i = 0; i < 2 * nstep; ++i) {
    E = dmrg_update(st, opt, log);
    st.oc += st.step;
    if (st.oc == 0 || st.oc == L - 2) st.step = -st.step;
  }
  return item(E);
}

double dmrg_run(DmrgState& st, const DmrgOptions& opt, DmrgLog* log, const DmrgHooks* hooks) {
  dmrg_prepare(st);
  const int L = int(st.mpo.size());
  // details::dmrg_impl (dmrg.cpp:224-236): sweeps start toward the far end
  // and come back; a centre on the last site starts one site to the left
  int init_pos = st.oc;
  st.step = st.oc == 0 ? 1 : -1;
  if (st.w12.size() == 1) st.step = 0;
  if (st.oc == L - 1) {
    --init_pos;
    --st.oc;
  }
  double E0 = 100000.0;
  if (hooks && hooks->init) hooks->init();
  size_t it = 0;
  for (; it < opt.maximum_iterations; ++it) {
    auto t0 = std::chrono::steady_clock::now();
    const double E = dmrg_sweep(st, opt, log);
    auto t1 = std::chrono::steady_clock::now();
    if (hooks && hooks->it_log_all) hooks->it_log_all(it, E, st.mps);
    if (log) {
      log->energies.push_back(E);
      log->sweep_ms.push_back(std::chrono::duration<double, std::milli>(t1 - t0).count());
      int64_t mb = 0;
      for (auto& a : st.mps) mb = std::max(mb, bond_of(a));
      log->max_bond.push_back(mb);
      log->middle_bond.push_back(bond_of(st.mps[st.mps.size() / 2]));
    }
    const double Eold = E0;
    E0 = E;
    if (!(std::fabs((E0 - Eold) / E0) > opt.convergence_criterion)) break;
  }
  // dmrg.cpp:257-266: back to where the centre started
  if (st.oc != init_pos) {
    if (st.oc != init_pos - 1 && init_pos != L - 1)
      throw Error(kRuntimeError, "the orthogonality center finished somewhere surprising! final oc: " +
                                     std::to_string(st.oc) + ". original oc: " + std::to_string(init_pos));
    mps_move_oc(st.mps, st.oc, init_pos);
  }
  if (hooks && hooks->end_log_all) hooks->end_log_all(it, E0, st.mps);
  return E0;
}

double contract_norm(const std::vector<BTensor>& a, const std::vector<BTensor>& b) {
  Context* ctx = a[0].ctx;
  Shape l1 = a.front().shape.shape_from({-1, 0, 0}).inverse_cvals();
  Shape l2 = b.front().shape.shape_from({-1, 0, 0});
  Shape r1 = a.back().shape.shape_from({0, 0, -1}).inverse_cvals();
  Shape r2 = b.back().shape.shape_from({0, 0, -1});
  BTensor left = eye_like(ctx, shape_product({&l1, &l2}));
  BTensor right = eye_like(ctx, shape_product({&r1, &r2}));
  for (size_t i = 0; i < a.size(); ++i) {
    left = tensordot(left, a[i], {0}, {0});
    left = tensordot(left, conj(b[i]), {0, 1}, {0, 1});
  }
  return item(tensordot(left, right, {0, 1}, {0, 1}));
}

double contract_energy(const std::vector<BTensor>& a, const std::vector<BTensor>& b, const std::vector<BTensor>& mpo) {
  Context* ctx = a[0].ctx;
  Shape l1 = a.front().shape.shape_from({-1, 0, 0}).inverse_cvals();
  Shape l2 = b.front().shape.shape_from({-1, 0, 0});
  Shape r1 = a.back().shape.shape_from({0, 0, -1}).inverse_cvals();
  Shape r2 = b.back().shape.shape_from({0, 0, -1});
  BTensor left = eye_like(ctx, shape_product({&l1, &l2}));
  BTensor right = eye_like(ctx, shape_product({&r1, &r2}));
  left = permute(tensordot(left, fill_like(ctx, edge_shape_prep(mpo.front().shape, 0), 1.0), {}, {}), {0, 2, 1});
  right = permute(tensordot(right, fill_like(ctx, edge_shape_prep(mpo.back().shape, 2), 1.0), {}, {}), {0, 2, 1});
  for (size_t i = 0; i < a.size(); ++i) {
    left = tensordot(left, a[i], {0}, {0});
    left = tensordot(left, mpo[i], {0, 2}, {0, 3});
    left = tensordot(left, conj(b[i]), {0, 2}, {0, 1});
  }
  return item(tensordot(left, right, {0, 1, 2}, {0, 1, 2}));
}

void mps_move_oc(std::vector<BTensor>& mps, int& oc, int target) {
  const int L = int(mps.size());
  if (target < 0 || target >= L) throw InvalidArgument("Proposed orthogonality center falls outside the MPS");
  if (oc < 0 || oc >= L) throw InvalidArgument("orthogonality center outside the MPS");
  while (target < oc) {
    SvdResult r = svd(mps[oc], 1);
    mps[oc] = permute(conj(r.V), {2, 0, 1});
    mps[oc - 1] = tensordot(mps[oc - 1], mul(r.U, r.d), {2}, {0});
    --oc;
  }
  while (target > oc) {
    SvdResult r = svd(mps[oc], 2);
    mps[oc] = std::move(r.U);
    mps[oc + 1] = tensordot(conj(mul(r.V, r.d)), mps[oc + 1], {0}, {0});
    ++oc;
  }
}

void mps_prepare(std::vector<BTensor>& mps) {
  const int L = int(mps.size());
  if (L < 1) throw InvalidArgument("empty MPS");
  const double n2 = contract_norm(mps, mps);
  if (!(n2 > 0.0)) throw DomainError("MPS of zero norm");
  mps[0] = scale(mps[0], 1.0 / std::sqrt(n2));
  int oc = 0;
  mps_move_oc(mps, oc, L - 1);
  mps_move_oc(mps, oc, 0);
}

}  // namespace qtt
