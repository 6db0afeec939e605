// Host bookkeeping of block-sparse tensors: shapes, block lists and the
// structural transforms (reshape, reshape_as, permute, conj).
#include "btensor.hpp"

#include <algorithm>
#include <numeric>

#include "kernels.hpp"

namespace qtt {

std::string fmt_idx(const BIdx& b, int r) {
  std::string s = "[";
  for (int k = 0; k < r; ++k) {
    if (k) s += ",";
    s += std::to_string(b.i[k]);
  }
  return s + "]";
}

std::vector<int64_t> Shape::sizes() const {
  std::vector<int64_t> out(rank());
  for (int d = 0; d < rank(); ++d) out[d] = std::accumulate(secs[d].begin(), secs[d].end(), int64_t(0));
  return out;
}

Charge Shape::charge_sum(const BIdx& b) const {
  Charge c = q_neutral();
  for (int d = 0; d < rank(); ++d) c = q_add(g, c, cvals[d][b.i[d]]);
  return c;
}

int64_t Shape::block_numel(const BIdx& b) const {
  int64_t n = 1;
  for (int d = 0; d < rank(); ++d) n *= secs[d][b.i[d]];
  return n;
}

void Shape::block_dims(const BIdx& b, int64_t* out) const {
  for (int d = 0; d < rank(); ++d) out[d] = secs[d][b.i[d]];
}

void Shape::validate() const {
  if (rank() > kMaxRank) throw InvalidArgument("rank larger than " + std::to_string(kMaxRank));
  if (secs.size() != cvals.size()) throw InvalidArgument("sections and charges have different ranks");
  for (int d = 0; d < rank(); ++d) {
    if (secs[d].size() != cvals[d].size())
      throw InvalidArgument("dimension " + std::to_string(d) + ": section and charge counts differ");
    for (auto s : secs[d])
      if (s < 0) throw InvalidArgument("negative section size");
  }
  if (g.n < 1 || g.n > kMaxGroups) throw InvalidArgument("bad number of charge groups");
}

// member shape_from (sources/btensor.cpp:492-583): keep the dims marked -1;
// for a fixed dim, the inverse charge of the section holding the element is
// absorbed into the selection rule.
Shape Shape::shape_from(const std::vector<int64_t>& dims) const {
  if (int(dims.size()) != rank()) throw InvalidArgument("The argument lenght must match the tensor's rank");
  auto sz = sizes();
  for (int d = 0; d < rank(); ++d)
    if (!(dims[d] < sz[d] || dims[d] == -1))
      throw InvalidArgument("The tensor doesn't have that many element");
  Shape out;
  out.g = g;
  out.sel = sel;
  for (int d = 0; d < rank(); ++d) {
    if (dims[d] < 0) {
      out.secs.push_back(secs[d]);
      out.cvals.push_back(cvals[d]);
    } else {
      int64_t e = dims[d];
      int s = 0;
      while (s < nsec(d) && e >= secs[d][s]) {
        e -= secs[d][s];
        ++s;
      }
      out.sel = q_add(g, out.sel, q_inv(g, cvals[d][s]));
    }
  }
  return out;
}

Shape Shape::inverse_cvals() const {
  Shape o = *this;
  for (auto& cv : o.cvals)
    for (auto& c : cv) c = q_inv(g, c);
  o.sel = q_inv(g, o.sel);
  return o;
}

Shape& Shape::neutral_shape_() {
  sel = q_neutral();
  for (auto& cv : cvals)
    for (auto& c : cv) c = q_neutral();
  return *this;
}

Shape shape_product(const std::vector<const Shape*>& parts) {
  Shape out;
  out.g = parts.at(0)->g;
  out.sel = q_neutral();
  for (auto* p : parts) {
    if (p->g != out.g) throw CheckError("the two tensors have different type of conserved quantities");
    for (int d = 0; d < p->rank(); ++d) {
      out.secs.push_back(p->secs[d]);
      out.cvals.push_back(p->cvals[d]);
    }
    out.sel = q_add(out.g, out.sel, p->sel);
  }
  return out;
}

int64_t BTensor::numel() const {
  int64_t n = 0;
  for (auto& b : blocks) n += b.numel;
  return n;
}

const Block* BTensor::find(const BIdx& idx) const {
  const int r = rank();
  auto it = std::lower_bound(blocks.begin(), blocks.end(), idx,
                             [r](const Block& b, const BIdx& k) { return bidx_less(b.idx, k, r); });
  if (it != blocks.end() && bidx_eq(it->idx, idx, r)) return &*it;
  return nullptr;
}

void BTensor::allocate(const std::vector<BIdx>& idx) {
  const int r = rank();
  // indices are checked before any section size is read through them
  for (auto& i : idx)
    for (int d = 0; d < r; ++d)
      if (i.i[d] < 0 || i.i[d] >= shape.nsec(d)) throw InvalidArgument("block index " + fmt_idx(i, r) + " out of range");
  blocks.clear();
  blocks.reserve(idx.size());
  int64_t off = 0;
  for (auto& i : idx) {
    int64_t n = shape.block_numel(i);
    blocks.push_back({i, off, n});
    off += n;
  }
  storage = ctx->alloc(size_t(std::max<int64_t>(off, 1)) * sizeof(double));
}

void BTensor::check_blocks_allowed() const {
  const int r = rank();
  for (size_t k = 0; k < blocks.size(); ++k) {
    const auto& b = blocks[k];
    for (int d = 0; d < r; ++d)
      if (b.idx.i[d] < 0 || b.idx.i[d] >= shape.nsec(d))
        throw InvalidArgument("block index " + fmt_idx(b.idx, r) + " out of range");
    if (!shape.allowed(b.idx))
      throw InvalidArgument("block index " + fmt_idx(b.idx, r) + " not allowed by selection rule");
    if (k && !bidx_less(blocks[k - 1].idx, b.idx, r))
      throw InvalidArgument("block list must be sorted and unique");
  }
}

void BTensor::upload(const std::vector<BIdx>& idx, const double* host) {
  allocate(idx);
  check_blocks_allowed();
  const int64_t n = numel();
  if (n) QTT_HIP(hipMemcpyAsync(base(), host, size_t(n) * sizeof(double), hipMemcpyHostToDevice, ctx->stream()));
  ctx->sync();
}

std::vector<double> BTensor::download() const {
  std::vector<double> out(static_cast<size_t>(numel()));
  if (!out.empty()) {
    // blocks may be a subset of the storage (views); copy block by block
    size_t pos = 0;
    for (auto& b : blocks) {
      QTT_HIP(hipMemcpyAsync(out.data() + pos, ptr(b), size_t(b.numel) * sizeof(double), hipMemcpyDeviceToHost,
                             ctx->stream()));
      pos += size_t(b.numel);
    }
    ctx->sync();
  }
  return out;
}

// ---------------------------------------------------------------------------
// reshape (sources/btensor.cpp:2986-3024): legs [g_k, g_{k+1}) fuse into one
// leg enumerating their section combinations row-major (last index fastest);
// fused sizes multiply and charges add; block data keeps its row-major layout.
// ---------------------------------------------------------------------------
BTensor reshape(const BTensor& t, const std::vector<int64_t>& index_groups) {
  const Shape& s = t.shape;
  std::vector<int64_t> bounds;
  bounds.push_back(0);
  for (auto x : index_groups) bounds.push_back(x);
  bounds.push_back(s.rank());
  for (size_t i = 1; i < bounds.size(); ++i)
    if (bounds[i] < bounds[i - 1] || bounds[i] > s.rank()) throw InvalidArgument("invalid index groups for reshape");
  Shape o;
  o.g = s.g;
  o.sel = s.sel;
  const int orank = int(bounds.size()) - 1;
  if (orank > kMaxRank) throw InvalidArgument("reshape rank too large");
  for (int gi = 0; gi < orank; ++gi) {
    std::vector<int64_t> sz{1};
    std::vector<Charge> ch{q_neutral()};
    for (int64_t d = bounds[gi]; d < bounds[gi + 1]; ++d) {
      std::vector<int64_t> nsz;
      std::vector<Charge> nch;
      for (size_t a = 0; a < sz.size(); ++a)
        for (int b = 0; b < s.nsec(int(d)); ++b) {
          nsz.push_back(sz[a] * s.secs[d][b]);
          nch.push_back(q_add(s.g, ch[a], s.cvals[d][b]));
        }
      sz.swap(nsz);
      ch.swap(nch);
    }
    o.secs.push_back(sz);
    o.cvals.push_back(ch);
  }
  BTensor out(t.ctx, o);
  out.storage = t.storage;
  out.blocks.reserve(t.blocks.size());
  for (auto& b : t.blocks) {
    Block nb{BIdx{}, b.off, b.numel};
    for (int gi = 0; gi < orank; ++gi) {
      int64_t flat = 0;
      for (int64_t d = bounds[gi]; d < bounds[gi + 1]; ++d) flat = flat * s.nsec(int(d)) + b.idx.i[d];
      nb.idx.i[gi] = int32_t(flat);
    }
    out.blocks.push_back(nb);
  }
  return out;
}

// reshape_as (sources/btensor.cpp:3026-3083)
BTensor reshape_as(const BTensor& t, const Shape& other, bool overwrite_c_vals) {
  const Shape& s = t.shape;
  int64_t n_self = 1, n_other = 1;
  for (int d = 0; d < s.rank(); ++d) n_self *= s.nsec(d);
  for (int d = 0; d < other.rank(); ++d) n_other *= other.nsec(d);
  if (n_self != n_other) throw InvalidArgument("incompatible sections layouts");
  // every possible block: same flux (dims_only) and same element count
  {
    BIdx a{}, b{};
    for (int64_t f = 0; f < n_self; ++f) {
      if (!overwrite_c_vals && s.charge_sum(a) != other.charge_sum(b))
        throw InvalidArgument("incompatible conserved quantities");
      if (s.block_numel(a) != other.block_numel(b)) throw InvalidArgument("incompatible block dimensions");
      for (int d = s.rank() - 1; d >= 0; --d) {
        if (++a.i[d] < s.nsec(d)) break;
        a.i[d] = 0;
      }
      for (int d = other.rank() - 1; d >= 0; --d) {
        if (++b.i[d] < other.nsec(d)) break;
        b.i[d] = 0;
      }
    }
  }
  Shape o = other;
  if (!overwrite_c_vals) o.sel = s.sel;
  BTensor out(t.ctx, o);
  out.storage = t.storage;
  for (auto& b : t.blocks) {
    int64_t flat = 0;
    for (int d = 0; d < s.rank(); ++d) flat = flat * s.nsec(d) + b.idx.i[d];
    Block nb{BIdx{}, b.off, b.numel};
    for (int d = other.rank() - 1; d >= 0; --d) {
      nb.idx.i[d] = int32_t(flat % other.nsec(d));
      flat /= other.nsec(d);
    }
    if (overwrite_c_vals && !other.allowed(nb.idx))
      throw InvalidArgument("block at " + fmt_idx(b.idx, s.rank()) +
                            " in orginal btensor not allowed by the new selection rule");
    out.blocks.push_back(nb);
  }
  return out;
}

BTensor inverse_cvals(const BTensor& t) {
  BTensor out = t;
  out.shape = t.shape.inverse_cvals();
  return out;
}

BTensor conj(const BTensor& t) { return inverse_cvals(t); }

static void chunk_tasks(std::vector<kern::ChunkTask>& tasks, int desc, int64_t numel) {
  const int64_t nch = (numel + kern::kChunk - 1) / kern::kChunk;
  for (int64_t c = 0; c < nch; ++c) tasks.push_back({desc, int32_t(c)});
}

BTensor clone(const BTensor& t) {
  BTensor out(t.ctx, t.shape);
  std::vector<BIdx> idx;
  for (auto& b : t.blocks) idx.push_back(b.idx);
  out.allocate(idx);
  std::vector<kern::Copy2D> cp;
  std::vector<kern::ChunkTask> tasks;
  double* ab = reinterpret_cast<double*>(t.ctx->arena().base());
  for (size_t k = 0; k < t.blocks.size(); ++k) {
    const auto& b = t.blocks[k];
    if (!b.numel) continue;
    kern::Copy2D c{};
    c.src_off = t.ptr(b) - ab;
    c.dst_off = out.ptr(out.blocks[k]) - ab;
    c.rows = 1;
    c.cols = int32_t(b.numel);
    c.src_rs = c.dst_rs = b.numel;
    c.src_cs = c.dst_cs = 1;
    c.alpha = 1.0;
    chunk_tasks(tasks, int(cp.size()), b.numel);
    cp.push_back(c);
  }
  if (!tasks.empty()) {
    KernelTimer kt(*t.ctx, "copy2d");
    auto [dd, dt] = t.ctx->upload_pack(cp, tasks);
    kern::copy2d(t.ctx->stream(), ab, dd, dt, int(tasks.size()));
  }
  return out;
}

// permute (sources/btensor.cpp:1754-1802): block indices permuted and the
// list re-sorted; block data copied into the permuted row-major layout.
BTensor permute(const BTensor& t, const std::vector<int64_t>& in_perm) {
  const int r = t.rank();
  if (int(in_perm.size()) != r) throw InvalidArgument("permutation length must match the rank");
  std::vector<int64_t> perm(r);
  std::vector<int> seen(r, 0);
  for (int i = 0; i < r; ++i) {
    perm[i] = in_perm[i] < 0 ? in_perm[i] + r : in_perm[i];
    if (perm[i] < 0 || perm[i] >= r || seen[perm[i]]++) throw InvalidArgument("invalid permutation");
  }
  Shape o;
  o.g = t.shape.g;
  o.sel = t.shape.sel;
  for (int i = 0; i < r; ++i) {
    o.secs.push_back(t.shape.secs[perm[i]]);
    o.cvals.push_back(t.shape.cvals[perm[i]]);
  }
  BTensor out(t.ctx, o);
  struct Ent {
    BIdx idx;
    size_t src;
  };
  std::vector<Ent> ents;
  ents.reserve(t.blocks.size());
  for (size_t k = 0; k < t.blocks.size(); ++k) {
    BIdx ni{};
    for (int i = 0; i < r; ++i) ni.i[i] = t.blocks[k].idx.i[perm[i]];
    ents.push_back({ni, k});
  }
  std::sort(ents.begin(), ents.end(), [r](const Ent& a, const Ent& b) { return bidx_less(a.idx, b.idx, r); });
  std::vector<BIdx> idx;
  for (auto& e : ents) idx.push_back(e.idx);
  out.allocate(idx);
  bool identity = true;
  for (int i = 0; i < r; ++i) identity &= perm[i] == i;
  double* ab = reinterpret_cast<double*>(t.ctx->arena().base());
  std::vector<kern::PermDesc> descs;
  std::vector<kern::ChunkTask> tasks;
  for (size_t k = 0; k < ents.size(); ++k) {
    const Block& sb = t.blocks[ents[k].src];
    if (!sb.numel) continue;
    int64_t dims[kMaxRank], st[kMaxRank];
    t.shape.block_dims(sb.idx, dims);
    st[r - 1] = 1;
    for (int i = r - 2; i >= 0; --i) st[i] = st[i + 1] * dims[i + 1];
    kern::PermDesc d{};
    d.src_off = t.ptr(sb) - ab;
    d.dst_off = out.ptr(out.blocks[k]) - ab;
    d.numel = sb.numel;
    if (identity) {
      d.rank = 1;
      d.dims[0] = int32_t(sb.numel);
      d.sstride[0] = 1;
    } else {
      d.rank = r;
      for (int i = 0; i < r; ++i) {
        d.dims[i] = int32_t(dims[perm[i]]);
        d.sstride[i] = st[perm[i]];
      }
    }
    chunk_tasks(tasks, int(descs.size()), sb.numel);
    descs.push_back(d);
  }
  if (!tasks.empty()) {
    KernelTimer kt(*t.ctx, "permute_copy");
    auto [dd, dt] = t.ctx->upload_pack(descs, tasks);
    kern::permute_copy(t.ctx->stream(), ab, dd, dt, int(tasks.size()));
  }
  return out;
}

}  // namespace qtt
