// Elementwise operations on block-sparse tensors: block matching on the host
// (the reference's broadcast / merge semantics), arithmetic on device.
#include <algorithm>
#include <cmath>

#include "btensor.hpp"
#include "kernels.hpp"

namespace qtt {

namespace {

double* arena_base(Context* c) { return reinterpret_cast<double*>(c->arena().base()); }

void add_chunks(std::vector<kern::ChunkTask>& tasks, int desc, int64_t numel) {
  const int64_t nch = (numel + kern::kChunk - 1) / kern::kChunk;
  for (int64_t c = 0; c < nch; ++c) tasks.push_back({desc, int32_t(c)});
}

void require_same_layout(const Shape& x, const Shape& y) {
  // add_tensor_check (sources/btensor.cpp:2320-2328)
  if (x.g != y.g || x.cvals != y.cvals)
    throw InvalidArgument("The conserved quantities of the tensors must be a perfect match");
  if (x.sel != y.sel) throw InvalidArgument("The selection rules of the tensors must be the same");
  if (x.secs != y.secs) throw InvalidArgument("The blocks of the tensors must have the same dimensions");
}

}  // namespace

// Broadcast multiply (btensor.cpp:1631-1689 with mul_helpers :977-1180) for
// the shapes the DMRG path uses: a rank-0 factor, a rank-1 factor along the
// last dimension, or equal shapes.  Only block pairs whose trailing indices
// match produce an output block.
BTensor mul(const BTensor& x_in, const BTensor& y_in) {
  const bool swap = x_in.rank() < y_in.rank();
  const BTensor& x = swap ? y_in : x_in;  // larger rank
  const BTensor& y = swap ? x_in : y_in;
  Context* ctx = x.ctx;
  const int lr = x.rank(), sr = y.rank();
  Shape o = x.shape;
  // shape_compute: trailing dims take the smaller tensor's sections when the
  // counts are equal; charges multiply (group op)
  for (int n = 1; n <= sr; ++n) {
    const int dl = lr - n, ds = sr - n;
    const bool s_one = y.shape.nsec(ds) == 1 && y.shape.secs[ds][0] == 1;
    const bool l_one = x.shape.nsec(dl) == 1 && x.shape.secs[dl][0] == 1;
    // both singleton: the sections coincide and the product is elementwise
    if (s_one != l_one) throw InvalidArgument("broadcast over singleton dimensions is not supported");
    if (x.shape.nsec(dl) != y.shape.nsec(ds))
      throw InvalidArgument("The section and tensor size of tensor a must match the section size and tensor size of "
                            "tensor b at non-singleton dimension " +
                            std::to_string(dl));
    o.secs[dl] = y.shape.secs[ds];
    for (int s = 0; s < x.shape.nsec(dl); ++s) o.cvals[dl][s] = q_add(o.g, y.shape.cvals[ds][s], x.shape.cvals[dl][s]);
  }
  o.sel = q_add(o.g, x.shape.sel, y.shape.sel);
  BTensor out(ctx, o);
  struct P {
    int xi, yi;
  };
  std::vector<P> pairs;
  if (sr == 0) {
    if (!y.blocks.empty())
      for (size_t k = 0; k < x.blocks.size(); ++k) pairs.push_back({int(k), 0});
  } else {
    for (size_t k = 0; k < x.blocks.size(); ++k) {
      BIdx key{};
      for (int j = 0; j < sr; ++j) key.i[j] = x.blocks[k].idx.i[lr - sr + j];
      const Block* yb = y.find(key);
      if (yb) pairs.push_back({int(k), int(yb - y.blocks.data())});
    }
  }
  std::vector<BIdx> idx;
  for (auto& p : pairs) idx.push_back(x.blocks[p.xi].idx);
  out.allocate(idx);
  double* ab = arena_base(ctx);
  std::vector<kern::EwDesc> descs;
  std::vector<kern::ChunkTask> tasks;
  for (size_t k = 0; k < pairs.size(); ++k) {
    const Block& xb = x.blocks[pairs[k].xi];
    const Block& yb = y.blocks[pairs[k].yi];
    if (xb.numel != out.blocks[k].numel) throw InvalidArgument("broadcast block size mismatch");
    if (!xb.numel) continue;
    kern::EwDesc d{};
    d.x_off = x.ptr(xb) - ab;
    d.y_off = y.ptr(yb) - ab;
    d.out_off = out.ptr(out.blocks[k]) - ab;
    d.numel = xb.numel;
    if (sr == 0) {
      d.mode = 0;
    } else if (sr == 1 && lr > 1) {
      d.mode = 1;
      d.ncol = int32_t(yb.numel);
      if (yb.numel == 0 || xb.numel % yb.numel) throw InvalidArgument("broadcast block size mismatch");
    } else {
      if (sr != lr) throw InvalidArgument("unsupported broadcast");
      d.mode = 2;
    }
    add_chunks(tasks, int(descs.size()), xb.numel);
    descs.push_back(d);
  }
  if (!tasks.empty()) {
    KernelTimer kt(*ctx, "ew_mul");
    auto [up_descs, up_tasks] = ctx->upload_pack(descs, tasks);
    kern::ew_binary(ctx->stream(), ab, up_descs, up_tasks, int(tasks.size()), 0, 0.0);
  }
  return out;
}

// x / s for a rank-0 s (broadcast div_), skipped when |s| < min_abs: the
// `non_singular` guard of one_step_lanczos (sources/dmrg.cpp:597-603)
// evaluated on device.
BTensor div_scalar_tensor(const BTensor& x, const BTensor& s, double min_abs) {
  if (s.rank() != 0) throw InvalidArgument("divisor must be rank 0");
  Context* ctx = x.ctx;
  BTensor out(ctx, x.shape);
  std::vector<BIdx> idx;
  for (auto& b : x.blocks) idx.push_back(b.idx);
  out.allocate(idx);
  if (s.blocks.empty()) {
    // the reference's item() of an empty rank-0 tensor is 0 < min_abs: no division
    return clone(x);
  }
  double* ab = arena_base(ctx);
  std::vector<kern::EwDesc> descs;
  std::vector<kern::ChunkTask> tasks;
  for (size_t k = 0; k < x.blocks.size(); ++k) {
    if (!x.blocks[k].numel) continue;
    kern::EwDesc d{};
    d.x_off = x.ptr(x.blocks[k]) - ab;
    d.y_off = s.ptr(s.blocks[0]) - ab;
    d.out_off = out.ptr(out.blocks[k]) - ab;
    d.numel = x.blocks[k].numel;
    d.mode = 0;
    add_chunks(tasks, int(descs.size()), d.numel);
    descs.push_back(d);
  }
  if (!tasks.empty()) {
    KernelTimer kt(*ctx, "ew_div");
    auto [up_descs, up_tasks] = ctx->upload_pack(descs, tasks);
    kern::ew_binary(ctx->stream(), ab, up_descs, up_tasks, int(tasks.size()), 1,
                    min_abs);
  }
  return out;
}

namespace {

// union of two sorted block lists (flat_map::merge, include/blockTensor/flat_map.h:602-630)
BTensor merge_axpby(const BTensor& x, int64_t a_off, double a, const BTensor& y, int64_t b_off, double b,
                    bool x_present, bool y_present) {
  Context* ctx = x.ctx;
  require_same_layout(x.shape, y.shape);
  const int r = x.rank();
  BTensor out(ctx, x.shape);
  struct Ent {
    BIdx idx;
    int xi, yi;
  };
  std::vector<Ent> ents;
  size_t i = 0, j = 0;
  const size_t nx = x_present ? x.blocks.size() : 0, ny = y_present ? y.blocks.size() : 0;
  while (i < nx || j < ny) {
    if (j >= ny || (i < nx && bidx_less(x.blocks[i].idx, y.blocks[j].idx, r))) {
      ents.push_back({x.blocks[i].idx, int(i), -1});
      ++i;
    } else if (i >= nx || bidx_less(y.blocks[j].idx, x.blocks[i].idx, r)) {
      ents.push_back({y.blocks[j].idx, -1, int(j)});
      ++j;
    } else {
      ents.push_back({x.blocks[i].idx, int(i), int(j)});
      ++i;
      ++j;
    }
  }
  std::vector<BIdx> idx;
  for (auto& e : ents) idx.push_back(e.idx);
  out.allocate(idx);
  double* ab = arena_base(ctx);
  std::vector<kern::AxpbyDesc> descs;
  std::vector<kern::ChunkTask> tasks;
  for (size_t k = 0; k < ents.size(); ++k) {
    const int64_t n = out.blocks[k].numel;
    if (!n) continue;
    kern::AxpbyDesc d{};
    d.x_off = ents[k].xi >= 0 ? x.ptr(x.blocks[ents[k].xi]) - ab : -1;
    d.y_off = ents[k].yi >= 0 ? y.ptr(y.blocks[ents[k].yi]) - ab : -1;
    d.out_off = out.ptr(out.blocks[k]) - ab;
    d.numel = n;
    add_chunks(tasks, int(descs.size()), n);
    descs.push_back(d);
  }
  if (!tasks.empty()) {
    KernelTimer kt(*ctx, "axpby");
    auto [up_descs, up_tasks] = ctx->upload_pack(descs, tasks);
    kern::axpby(ctx->stream(), ab, up_descs, up_tasks, int(tasks.size()), a_off, a,
                b_off, b);
  }
  return out;
}

}  // namespace

// add(other, alpha) (sources/btensor.cpp:2666-2683): collisions a + alpha*b,
// blocks only in y inserted as alpha*b.
BTensor add(const BTensor& x, const BTensor& y, double alpha) {
  return merge_axpby(x, -1, 1.0, y, -1, alpha, true, true);
}

// a*x + b*y for rank-0 device scalars a and b: o*state + n*psi
// (sources/dmrg.cpp:635).  A rank-0 factor without a block is zero and the
// product then has no blocks (broadcast semantics).
BTensor axpby_dev(const BTensor& x, const BTensor& a, const BTensor& y, const BTensor& b) {
  if (a.rank() != 0 || b.rank() != 0) throw InvalidArgument("coefficients must be rank 0");
  double* ab = arena_base(x.ctx);
  const int64_t ao = a.blocks.empty() ? -1 : a.ptr(a.blocks[0]) - ab;
  const int64_t bo = b.blocks.empty() ? -1 : b.ptr(b.blocks[0]) - ab;
  return merge_axpby(x, ao, 0.0, y, bo, 0.0, !a.blocks.empty(), !b.blocks.empty());
}

BTensor scale(const BTensor& x, double alpha) {
  return merge_axpby(x, -1, alpha, x, -1, 0.0, true, false);
}

// Full contraction of all dims with y (real data): rank-0 result holding the
// sum over blocks present in both of sum(x*y).  No matching block -> no block.
BTensor full_dot(const BTensor& x, const BTensor& y, bool sqrt_out) {
  Context* ctx = x.ctx;
  if (x.rank() != y.rank()) throw CheckError("both dimension lists should have the same length.");
  std::vector<int64_t> dims(x.rank());
  for (int i = 0; i < x.rank(); ++i) dims[i] = i;
  check_product_compat(x.shape, inverse_cvals(y).shape, dims, dims, true);
  Shape o;
  o.g = x.shape.g;
  o.sel = q_add(o.g, x.shape.sel, q_inv(o.g, y.shape.sel));
  BTensor out(ctx, o);
  double* ab = arena_base(ctx);
  std::vector<kern::DotDesc> descs;
  std::vector<kern::ChunkTask> tasks;
  const int r = x.rank();
  size_t i = 0, j = 0;
  bool any = false;
  while (i < x.blocks.size() && j < y.blocks.size()) {
    if (bidx_less(x.blocks[i].idx, y.blocks[j].idx, r)) {
      ++i;
    } else if (bidx_less(y.blocks[j].idx, x.blocks[i].idx, r)) {
      ++j;
    } else {
      any = true;
      if (x.blocks[i].numel) {
        kern::DotDesc d{x.ptr(x.blocks[i]) - ab, y.ptr(y.blocks[j]) - ab, x.blocks[i].numel};
        add_chunks(tasks, int(descs.size()), d.numel);
        descs.push_back(d);
      }
      ++i;
      ++j;
    }
  }
  if (!any) return out;
  out.allocate({BIdx{}});
  BufPtr partial = ctx->alloc(std::max<size_t>(tasks.size(), 1) * sizeof(double));
  KernelTimer kt(*ctx, "dot");
  kern::dot(ctx->stream(), ab, tasks.empty() ? nullptr : ctx->upload_vec(descs),
            tasks.empty() ? nullptr : ctx->upload_vec(tasks), int(tasks.size()), partial->dptr() - ab,
            out.ptr(out.blocks[0]) - ab, sqrt_out);
  return out;
}

// ones_like / full_like: every block allowed by the selection rule, filled.
BTensor fill_like(Context* ctx, const Shape& s, double value) {
  BTensor out(ctx, s);
  std::vector<BIdx> idx;
  const int r = s.rank();
  BIdx b{};
  int64_t total = 1;
  for (int d = 0; d < r; ++d) total *= s.nsec(d);
  for (int64_t f = 0; f < total; ++f) {
    if (s.allowed(b)) idx.push_back(b);
    for (int d = r - 1; d >= 0; --d) {
      if (++b.i[d] < s.nsec(d)) break;
      b.i[d] = 0;
    }
  }
  out.allocate(idx);
  const int64_t n = out.numel();
  if (n) {
    double* ab = arena_base(ctx);
    kern::Copy2D c{};
    c.src_off = -1;
    c.dst_off = out.base() - ab;
    c.rows = 1;
    c.cols = int32_t(n);
    c.dst_cs = 1;
    c.alpha = value;
    std::vector<kern::Copy2D> cp{c};
    std::vector<kern::ChunkTask> tasks;
    add_chunks(tasks, 0, n);
    KernelTimer kt(*ctx, "fill");
    auto [up_cp, up_tasks] = ctx->upload_pack(cp, tasks);
    kern::copy2d(ctx->stream(), ab, up_cp, up_tasks, int(tasks.size()));
  }
  return out;
}

double item(const BTensor& s) {
  if (s.blocks.empty()) return 0.0;
  if (s.numel() != 1) throw LogicError("Only simgle block and single element tensors can be converted to scalar");
  double v = 0;
  s.ctx->read_small(&v, s.ptr(s.blocks[0]), sizeof(double));
  return v;
}

void normalize_d(BTensor& d) {
  // d blocks are contiguous in the storage of a freshly truncated d
  if (d.blocks.empty()) return;
  double* ab = arena_base(d.ctx);
  const int64_t first = d.blocks.front().off;
  int64_t expect = first;
  for (auto& b : d.blocks) {
    if (b.off != expect) throw LogicError("normalize_d expects contiguous blocks");
    expect += b.numel;
  }
  KernelTimer kt(*d.ctx, "normalize");
  kern::normalize(d.ctx->stream(), ab, d.ptr(d.blocks.front()) - ab, expect - first);
}

}  // namespace qtt
