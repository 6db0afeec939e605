// Block-sparse contraction: host-side sector matching that reproduces the
// output block list and accumulation order of btensor::tensordot
// (sources/btensor.cpp:1971-2119), and one grouped-GEMM launch on device.
#include <algorithm>
#include <climits>
#include <cmath>
#include <map>
#include <mutex>
#include <string>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <numeric>
#include <tuple>

#include "btensor.hpp"
#include "kernels.hpp"

namespace qtt {

// check_product_compat<Throws> (sources/btensor.cpp:782-825)
bool check_product_compat(const Shape& a, const Shape& b, const std::vector<int64_t>& da,
                          const std::vector<int64_t>& db, bool throws) {
  if (da.size() != db.size()) {
    if (throws) throw CheckError("both dimension lists should have the same length.");
    return false;
  }
  for (size_t i = 0; i < da.size(); ++i) {
    const int x = int(da[i]), y = int(db[i]);
    if (a.nsec(x) != b.nsec(y)) {
      if (throws)
        throw CheckError("contracted dimensions need to match, but first has " + std::to_string(a.nsec(x)) +
                         " sections along dim " + std::to_string(x) + "  and second has " +
                         std::to_string(b.nsec(y)) + " sections along dim " + std::to_string(y));
      return false;
    }
    for (int s = 0; s < a.nsec(x); ++s) {
      if (q_add(a.g, a.cvals[x][s], b.cvals[y][s]) != q_neutral()) {
        if (throws)
          throw CheckError("contracted conserved numbers need to sum to zero, but there is a violation when "
                           "contracting dim " +
                           std::to_string(x) + " of the left tensor with dim " + std::to_string(y) +
                           " of the right tensor");
        return false;
      }
    }
  }
  return true;
}

namespace {

// Element strides of a block's matrix view over the listed dims (in order);
// size-1 dims are skipped.  Returns false when the dims do not collapse into
// a single stride (the block then needs a packing copy).
bool collapse(const int64_t* dims, const int64_t* st, const std::vector<int>& which, int64_t& stride) {
  int last = -1;
  for (int d : which) {
    if (dims[d] == 1) continue;
    if (last >= 0 && st[last] != st[d] * dims[d]) return false;
    last = d;
  }
  stride = last >= 0 ? st[last] : 1;
  return true;
}

void add_chunks(std::vector<kern::ChunkTask>& tasks, int desc, int64_t numel) {
  const int64_t nch = (numel + kern::kChunk - 1) / kern::kChunk;
  for (int64_t c = 0; c < nch; ++c) tasks.push_back({desc, int32_t(c)});
}

struct KeyEnt {
  BIdx key;
  int src;
};

// ---- GEMM tile-variant tuning table (process-wide) -----------------------------
// A contraction's shape class: log2 of the work-weighted mean M and N of its
// MFMA outputs and of its 64 x 64 tile count.  Classes with fewer than
// kTuneMinTiles tiles always take the 64 x 64 tile (latency-bound).
constexpr int64_t kTuneMinTiles = 64;

struct TuneTable {
  std::mutex mu;
  std::map<std::string, int> best;
};
TuneTable& tune_table() {
  static TuneTable t;
  return t;
}

int ilog2(double x) { return x < 1.0 ? 0 : int(std::floor(std::log2(x))); }

std::string gemm_shape_key(const std::vector<kern::GemmOut>& outs, const std::vector<int64_t>& work, int64_t tiles) {
  double wm = 0, wn = 0, w = 0;
  for (size_t o = 0; o < outs.size(); ++o) {
    const double f = double(outs[o].M) * outs[o].N * double(std::max<int64_t>(work[o], 1));
    wm += f * outs[o].M;
    wn += f * outs[o].N;
    w += f;
  }
  if (w > 0) {
    wm /= w;
    wn /= w;
  }
  return "f64 m" + std::to_string(ilog2(wm)) + " n" + std::to_string(ilog2(wn)) + " t" +
         std::to_string(ilog2(double(tiles)));
}

// Variant for a class.  QTT_GEMM_VARIANT=<n> forces one; QTT_GEMM_TUNE=0
// disables tuning (64 x 64 unless the table has an entry).  *tune is set
// when the caller should time the variants now (unknown class, tuning on,
// timing allowed).
int gemm_variant_for(const std::string& key, int64_t tiles, bool may_time, bool* tune) {
  *tune = false;
  if (const char* f = std::getenv("QTT_GEMM_VARIANT")) {
    const int v = std::atoi(f);
    if (v >= 0 && v < kern::kGemmVariants) return v;
  }
  if (tiles < kTuneMinTiles) return kern::kTile64;
  {
    TuneTable& t = tune_table();
    std::lock_guard<std::mutex> lk(t.mu);
    auto it = t.best.find(key);
    if (it != t.best.end()) return it->second;
  }
  const char* e = std::getenv("QTT_GEMM_TUNE");
  *tune = may_time && !(e && std::atoi(e) == 0);
  return kern::kTile64;
}

void gemm_variant_record(const std::string& key, int v) {
  TuneTable& t = tune_table();
  std::lock_guard<std::mutex> lk(t.mu);
  t.best[key] = v;
}

// Where one contraction keeps its temporaries (packed operands, split-K
// partials) and its launch descriptors.
class TdWork {
 public:
  virtual ~TdWork() = default;
  // element offset (from the arena base) of `bytes` of device scratch
  virtual int64_t take(size_t bytes) = 0;
  // device copy of host bytes, ordered on the context stream before the
  // work enqueued afterwards
  virtual void* put(const void* src, size_t bytes) = 0;
  virtual bool timed() const = 0;
};

// Temporaries from the context arena, descriptors through its staging ring.
class ArenaWork : public TdWork {
 public:
  explicit ArenaWork(Context* c) : ctx_(c) {}
  int64_t take(size_t bytes) override {
    keep_.push_back(ctx_->alloc(bytes));
    return keep_.back()->dptr() - reinterpret_cast<double*>(ctx_->arena().base());
  }
  void* put(const void* src, size_t bytes) override { return ctx_->upload(src, bytes); }
  bool timed() const override { return true; }

 private:
  Context* ctx_;
  std::vector<BufPtr> keep_;  // released at the end of the call: later work is stream-ordered after it
};

// Everything in the device's static contraction scratch (no allocation):
// regions are handed out front to back; the call fails before any launch
// when they do not fit.
class StaticWork : public TdWork {
 public:
  explicit StaticWork(Context* c) : ctx_(c), base_(static_cast<char*>(c->noalloc_scratch())) {}
  int64_t take(size_t bytes) override {
    const size_t off = (pos_ + 255) / 256 * 256;
    if (off + bytes > kern::kNoAllocScratchBytes)
      throw InvalidArgument("contraction needs more than the " +
                            std::to_string(kern::kNoAllocScratchBytes >> 20) +
                            " MiB scratch of the allocation-free entry point (" + std::to_string(off + bytes) +
                            " bytes)");
    pos_ = off + bytes;
    return reinterpret_cast<double*>(base_ + off) - reinterpret_cast<double*>(ctx_->arena().base());
  }
  void* put(const void* src, size_t bytes) override {
    double* p = reinterpret_cast<double*>(ctx_->arena().base()) + take(bytes ? bytes : 1);
    ctx_->stage_to(p, src, bytes);
    return p;
  }
  bool timed() const override { return false; }

 private:
  Context* ctx_;
  char* base_;
  size_t pos_ = 0;
};

// Several descriptor lists in one copy; each list starts 16-byte aligned.
template <class... T>
std::tuple<T*...> put_pack(TdWork& w, const std::vector<T>&... v) {
  constexpr size_t kA = 16;
  const size_t sizes[] = {v.size() * sizeof(T)...};
  size_t offs[sizeof...(T)], total = 0;
  for (size_t i = 0; i < sizeof...(T); ++i) {
    offs[i] = total;
    total += (sizes[i] + kA - 1) / kA * kA;
  }
  std::vector<char> buf(total ? total : 1);
  const void* src[] = {static_cast<const void*>(v.data())...};
  for (size_t i = 0; i < sizeof...(T); ++i)
    if (sizes[i]) std::memcpy(buf.data() + offs[i], src[i], sizes[i]);
  char* d = static_cast<char*>(w.put(buf.data(), total));
  size_t i = 0;
  return std::tuple<T*...>{reinterpret_cast<T*>(d + offs[i++])...};
}

BTensor tensordot_impl(const BTensor& a, const BTensor& b, const std::vector<int64_t>& da_in,
                       const std::vector<int64_t>& db_in, ContractStats* stats, TdWork& work, const BTensor* into,
                       TdPlan::Launch* plan);

// Temporaries and descriptors in arena regions owned by a plan, so the plan
// can be launched again later (descriptors are copied to the device once).
class PersistWork : public TdWork {
 public:
  PersistWork(Context* c, std::vector<BufPtr>* keep) : ctx_(c), keep_(keep) {}
  int64_t take(size_t bytes) override {
    keep_->push_back(ctx_->alloc(bytes));
    return keep_->back()->dptr() - reinterpret_cast<double*>(ctx_->arena().base());
  }
  void* put(const void* src, size_t bytes) override {
    keep_->push_back(ctx_->alloc(bytes ? bytes : 1));
    ctx_->stage_to(keep_->back()->cptr(), src, bytes);
    return keep_->back()->cptr();
  }
  bool timed() const override { return false; }

 private:
  Context* ctx_;
  std::vector<BufPtr>* keep_;
};

}  // namespace

// Kernel arguments of one contraction: permute-packs, narrow-row tasks, MFMA
// tiles, split-K sums (device descriptor pointers).
struct TdPlan::Launch {
  const kern::PermDesc* pdesc = nullptr;
  const kern::ChunkTask* ptasks = nullptr;
  int nptasks = 0;
  const kern::RowTask* rows = nullptr;
  int nrows = 0;
  const kern::GemmTile* tiles = nullptr;
  int ntiles = 0;
  const kern::GemmOut* outs = nullptr;
  const kern::GemmPair* pairs = nullptr;
  const kern::SumParts* sums = nullptr;
  const kern::ChunkTask* sum_tasks = nullptr;
  int nsum_tasks = 0;
  bool split_fp32 = false;
  int variant = 0;  // FP64 tile variant (kern::GemmVariant)
  const char* gemm_name = "gemm_f64";
};

namespace {

void td_pack(Context* ctx, const TdPlan::Launch& l, bool timed) {
  if (!l.nptasks) return;
  std::unique_ptr<KernelTimer> kt;
  if (timed) kt = std::make_unique<KernelTimer>(*ctx, "pack_permute");
  kern::permute_copy(ctx->stream(), reinterpret_cast<double*>(ctx->arena().base()), l.pdesc, l.ptasks, l.nptasks);
}

void td_run(Context* ctx, const TdPlan::Launch& l, bool timed) {
  double* ab = reinterpret_cast<double*>(ctx->arena().base());
  auto timer = [&](const char* name) {
    return timed ? std::make_unique<KernelTimer>(*ctx, name) : std::unique_ptr<KernelTimer>();
  };
  td_pack(ctx, l, timed);
  if (l.nrows) {
    auto kt = timer("gemm_rows");
    kern::gemm_rows(ctx->stream(), ab, l.rows, l.nrows, l.outs, l.pairs, 1.0, 0.0);
  }
  if (l.split_fp32) {
    auto kt = timer(l.gemm_name);
    kern::gemm_split(ctx->stream(), ab, l.tiles, l.ntiles, l.outs, l.pairs, 1.0, 0.0);
  } else {
    auto kt = timer(l.gemm_name);
    kern::gemm_f64_variant(ctx->stream(), l.variant, ab, l.tiles, l.ntiles, l.outs, l.pairs, 1.0, 0.0);
  }
  if (l.nsum_tasks) {
    auto kt = timer("gemm_split_k_sum");
    kern::sum_parts(ctx->stream(), ab, l.sums, l.sum_tasks, l.nsum_tasks);
  }
}

}  // namespace

// The tuning table as JSON: {"gemm_variants": {"<class>": <variant>, ...}}.
void gemm_tuning_save(const std::string& path) {
  TuneTable& t = tune_table();
  std::lock_guard<std::mutex> lk(t.mu);
  std::FILE* f = std::fopen(path.c_str(), "w");
  if (!f) throw InvalidArgument("cannot write " + path);
  std::fprintf(f, "{\"gemm_variants\": {");
  bool first = true;
  for (auto& kv : t.best) {
    std::fprintf(f, "%s\n  \"%s\": %d", first ? "" : ",", kv.first.c_str(), kv.second);
    first = false;
  }
  std::fprintf(f, "\n}}\n");
  std::fclose(f);
}

void gemm_tuning_load(const std::string& path) {
  std::FILE* f = std::fopen(path.c_str(), "r");
  if (!f) throw InvalidArgument("cannot read " + path);
  std::string text;
  char buf[4096];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, n);
  std::fclose(f);
  const size_t start = text.find("\"gemm_variants\"");
  if (start == std::string::npos) throw InvalidArgument(path + ": no gemm_variants table");
  size_t pos = text.find('{', start);
  std::map<std::string, int> got;
  while (pos != std::string::npos) {
    const size_t q0 = text.find('"', pos + 1);
    const size_t close = text.find('}', pos + 1);
    if (q0 == std::string::npos || (close != std::string::npos && close < q0)) break;
    const size_t q1 = text.find('"', q0 + 1);
    const size_t colon = text.find(':', q1);
    if (q1 == std::string::npos || colon == std::string::npos) throw InvalidArgument(path + ": malformed table");
    const int v = std::atoi(text.c_str() + colon + 1);
    if (v < 0 || v >= kern::kGemmVariants) throw InvalidArgument(path + ": unknown variant");
    got[text.substr(q0 + 1, q1 - q0 - 1)] = v;
    pos = text.find_first_of(",}", colon);
    if (pos == std::string::npos || text[pos] == '}') break;
  }
  TuneTable& t = tune_table();
  std::lock_guard<std::mutex> lk(t.mu);
  for (auto& kv : got) t.best[kv.first] = kv.second;
}

size_t gemm_tuning_size() {
  TuneTable& t = tune_table();
  std::lock_guard<std::mutex> lk(t.mu);
  return t.best.size();
}

TdPlan plan_tensordot(const BTensor& a, const BTensor& b, const std::vector<int64_t>& da,
                      const std::vector<int64_t>& db, ContractStats* stats) {
  TdPlan p;
  p.launch = std::make_shared<TdPlan::Launch>();
  PersistWork w(a.ctx, &p.keep);
  p.out = tensordot_impl(a, b, da, db, stats, w, nullptr, p.launch.get());
  return p;
}

void run_tensordot(const TdPlan& p, bool timed) {
  if (p.launch && p.out.ctx) td_run(p.out.ctx, *p.launch, timed);
}

BTensor tensordot(const BTensor& a, const BTensor& b, const std::vector<int64_t>& da, const std::vector<int64_t>& db,
                  ContractStats* stats) {
  ArenaWork w(a.ctx);
  return tensordot_impl(a, b, da, db, stats, w, nullptr, nullptr);
}

void tensordot_into(const BTensor& a, const BTensor& b, const std::vector<int64_t>& da,
                    const std::vector<int64_t>& db, BTensor& out, ContractStats* stats) {
  if (out.ctx != a.ctx) throw InvalidArgument("tensors belong to different contexts");
  if (!out.storage && out.numel()) throw InvalidArgument("output tensor has no storage");
  // the output must not share memory with an operand
  if (out.storage && (out.storage == a.storage || out.storage == b.storage))
    throw InvalidArgument("output tensor shares storage with an operand");
  StaticWork w(a.ctx);
  tensordot_impl(a, b, da, db, stats, w, &out, nullptr);
}

namespace {

BTensor tensordot_impl(const BTensor& a, const BTensor& b, const std::vector<int64_t>& da_in,
                       const std::vector<int64_t>& db_in, ContractStats* stats, TdWork& work, const BTensor* into,
                       TdPlan::Launch* plan) {
  Context* ctx = a.ctx;
  TdPlan::Launch lch;
  HostTimer ht_match(*ctx, work.timed() ? "host_td_match" : nullptr);
  if (a.shape.g != b.shape.g) throw CheckError("the two tensors have different type of conserved quantities");
  if (da_in.size() != db_in.size()) throw CheckError("both dimension lists should have the same length.");
  const int ra = a.rank(), rb = b.rank();
  std::vector<int64_t> da(da_in), db(db_in);
  for (auto& d : da) {
    if (d < 0) d += ra;
    if (d < 0 || d >= ra) throw CheckError("dimension out of range");
  }
  for (auto& d : db) {
    if (d < 0) d += rb;
    if (d < 0 || d >= rb) throw CheckError("dimension out of range");
  }
  check_product_compat(a.shape, b.shape, da, db, true);
  const int dl = int(da.size());
  std::vector<int> free_a, free_b, p1, p2prime;
  for (int i = 0; i < ra; ++i)
    if (std::find(da.begin(), da.end(), i) == da.end()) free_a.push_back(i);
  for (int i = 0; i < rb; ++i)
    if (std::find(db.begin(), db.end(), i) == db.end()) free_b.push_back(i);
  p1 = free_a;
  for (auto d : da) p1.push_back(int(d));
  p2prime = free_b;
  for (auto d : db) p2prime.push_back(int(d));
  const int na = int(free_a.size()), nb = int(free_b.size());
  if (na + nb > kMaxRank) throw InvalidArgument("contraction result rank too large");

  Shape os;
  os.g = a.shape.g;
  for (int d : free_a) {
    os.secs.push_back(a.shape.secs[d]);
    os.cvals.push_back(a.shape.cvals[d]);
  }
  for (int d : free_b) {
    os.secs.push_back(b.shape.secs[d]);
    os.cvals.push_back(b.shape.cvals[d]);
  }
  os.sel = q_add(os.g, a.shape.sel, b.shape.sel);
  BTensor out(ctx, os);

  // permuted block keys (permute_bl, sources/btensor.cpp:1843-1894), sorted
  std::vector<KeyEnt> t1(a.blocks.size()), t2(b.blocks.size());
  for (size_t k = 0; k < a.blocks.size(); ++k) {
    t1[k].src = int(k);
    for (int j = 0; j < ra; ++j) t1[k].key.i[j] = a.blocks[k].idx.i[p1[j]];
  }
  for (size_t k = 0; k < b.blocks.size(); ++k) {
    t2[k].src = int(k);
    for (int j = 0; j < rb; ++j) t2[k].key.i[j] = b.blocks[k].idx.i[p2prime[j]];
  }
  std::sort(t1.begin(), t1.end(), [ra](const KeyEnt& x, const KeyEnt& y) { return bidx_less(x.key, y.key, ra); });
  std::sort(t2.begin(), t2.end(), [rb](const KeyEnt& x, const KeyEnt& y) { return bidx_less(x.key, y.key, rb); });

  auto prefix_eq = [](const BIdx& x, const BIdx& y, int n) {
    for (int j = 0; j < n; ++j)
      if (x.i[j] != y.i[j]) return false;
    return true;
  };
  // group index of every t2 entry (runs of equal free prefix)
  std::vector<int> g2of(t2.size());
  {
    int g = -1;
    for (size_t k = 0; k < t2.size(); ++k) {
      if (k == 0 || !prefix_eq(t2[k].key, t2[k - 1].key, nb)) ++g;
      g2of[k] = g;
    }
  }
  // t2 entries ordered by (contracted suffix, group): suffix lookup table
  std::vector<int> by_suffix(t2.size());
  std::iota(by_suffix.begin(), by_suffix.end(), 0);
  auto suffix_less2 = [&](int x, int y) {
    for (int j = 0; j < dl; ++j) {
      const int u = t2[x].key.i[nb + j], v = t2[y].key.i[nb + j];
      if (u != v) return u < v;
    }
    return g2of[x] < g2of[y];
  };
  std::sort(by_suffix.begin(), by_suffix.end(), suffix_less2);

  struct Match {
    int g2, e1, e2;
  };
  struct OutBlk {
    BIdx idx;
    int pb, pe;  // range in all_pairs
  };
  std::vector<OutBlk> outs;
  std::vector<std::pair<int, int>> all_pairs;  // (a block, b block), grouped by output block
  std::vector<Match> matches;
  size_t s1 = 0;
  while (s1 < t1.size()) {
    size_t e1 = s1 + 1;
    while (e1 < t1.size() && prefix_eq(t1[e1].key, t1[s1].key, na)) ++e1;
    matches.clear();
    for (size_t k = s1; k < e1; ++k) {
      // range of t2 entries whose suffix equals this entry's suffix
      auto cmp_lo = [&](int x) {
        for (int j = 0; j < dl; ++j) {
          const int u = t2[x].key.i[nb + j], v = t1[k].key.i[na + j];
          if (u != v) return u < v;
        }
        return false;
      };
      auto lo = std::partition_point(by_suffix.begin(), by_suffix.end(), cmp_lo);
      for (auto it = lo; it != by_suffix.end(); ++it) {
        bool eq = true;
        for (int j = 0; j < dl && eq; ++j) eq = t2[*it].key.i[nb + j] == t1[k].key.i[na + j];
        if (!eq) break;
        matches.push_back({g2of[*it], int(k), *it});
      }
    }
    std::stable_sort(matches.begin(), matches.end(), [](const Match& x, const Match& y) { return x.g2 < y.g2; });
    size_t m = 0;
    while (m < matches.size()) {
      size_t me = m + 1;
      while (me < matches.size() && matches[me].g2 == matches[m].g2) ++me;
      OutBlk ob;
      for (int j = 0; j < na; ++j) ob.idx.i[j] = t1[s1].key.i[j];
      for (int j = 0; j < nb; ++j) ob.idx.i[na + j] = t2[matches[m].e2].key.i[j];
      ob.pb = int(all_pairs.size());
      for (size_t q = m; q < me; ++q) all_pairs.push_back({t1[matches[q].e1].src, t2[matches[q].e2].src});
      ob.pe = int(all_pairs.size());
      outs.push_back(ob);
      m = me;
    }
    s1 = e1;
  }

  std::vector<BIdx> oidx;
  oidx.reserve(outs.size());
  for (auto& o : outs) oidx.push_back(o.idx);
  ht_match.stop();
  if (into) {
    // the caller's output: the structure this contraction produces, blocks
    // in the same order
    const Shape& is = into->shape;
    bool same = is.g == os.g && is.secs == os.secs && is.cvals == os.cvals && is.sel == os.sel &&
                into->blocks.size() == oidx.size();
    for (size_t k = 0; same && k < oidx.size(); ++k) same = bidx_eq(into->blocks[k].idx, oidx[k], na + nb);
    if (!same) throw InvalidArgument("output tensor does not have the structure of the contraction result");
    out = *into;
  } else {
    HostTimer ht(*ctx, "host_td_alloc");
    out.allocate(oidx);
  }
  if (outs.empty()) return out;
  HostTimer ht_plan(*ctx, work.timed() ? "host_td_plan" : nullptr);

  // matrix views of the operand blocks; blocks whose permuted view is not a
  // plain strided matrix are packed (row-major) into a scratch buffer
  double* ab = reinterpret_cast<double*>(ctx->arena().base());
  struct View {
    int64_t off = -1, rs = 0, cs = 0;
    bool need_pack = false;
    int64_t M = 0, K = 0;
  };
  std::vector<View> va(a.blocks.size()), vb(b.blocks.size());
  std::vector<char> used_a(a.blocks.size(), 0), used_b(b.blocks.size(), 0);
  for (auto& p : all_pairs) {
    used_a[p.first] = 1;
    used_b[p.second] = 1;
  }
  std::vector<int> ca(da.begin(), da.end()), cb(db.begin(), db.end());
  std::vector<kern::PermDesc> pdesc;
  std::vector<kern::ChunkTask> ptasks;
  int64_t pack_elems = 0;
  std::vector<std::pair<int64_t*, int64_t>> fixups;  // (pointer to off, element offset inside scratch)
  auto plan_view = [&](const BTensor& t, const Block& blk, const std::vector<int>& rows, const std::vector<int>& cols,
                       View& v) {
    const int r = t.rank();
    int64_t dims[kMaxRank], st[kMaxRank];
    t.shape.block_dims(blk.idx, dims);
    if (r) st[r - 1] = 1;
    for (int i = r - 2; i >= 0; --i) st[i] = st[i + 1] * dims[i + 1];
    v.M = 1;
    v.K = 1;
    for (int d : rows) v.M *= dims[d];
    for (int d : cols) v.K *= dims[d];
    int64_t rs, cs;
    if (collapse(dims, st, rows, rs) && collapse(dims, st, cols, cs)) {
      v.off = t.ptr(blk) - ab;
      v.rs = rs;
      v.cs = cs;
      // a unit-length side: keep the canonical contiguous stride
      if (v.M == 1) v.rs = v.K;
      if (v.K == 1) v.cs = 1;
      if (v.M == 1 && v.K > 1) v.cs = cs;
      return;
    }
    v.need_pack = true;
    kern::PermDesc d{};
    d.src_off = t.ptr(blk) - ab;
    d.dst_off = pack_elems;  // fixed up once the scratch is allocated
    d.numel = blk.numel;
    d.rank = int32_t(rows.size() + cols.size());
    int j = 0;
    for (int dd : rows) {
      d.dims[j] = int32_t(dims[dd]);
      d.sstride[j++] = st[dd];
    }
    for (int dd : cols) {
      d.dims[j] = int32_t(dims[dd]);
      d.sstride[j++] = st[dd];
    }
    v.off = pack_elems;
    v.rs = v.K;
    v.cs = 1;
    add_chunks(ptasks, int(pdesc.size()), blk.numel);
    pdesc.push_back(d);
    pack_elems += blk.numel;
  };
  for (size_t k = 0; k < a.blocks.size(); ++k)
    if (used_a[k]) plan_view(a, a.blocks[k], free_a, ca, va[k]);
  for (size_t k = 0; k < b.blocks.size(); ++k)
    if (used_b[k]) plan_view(b, b.blocks[k], cb, free_b, vb[k]);
  if (pack_elems) {
    const int64_t soff = work.take(size_t(pack_elems) * sizeof(double));
    for (auto& d : pdesc) d.dst_off += soff;
    for (auto& v : va)
      if (v.need_pack) v.off += soff;
    for (auto& v : vb)
      if (v.need_pack) v.off += soff;
    HostTimer ht(*ctx, work.timed() ? "host_td_upload" : nullptr);
    std::tie(lch.pdesc, lch.ptasks) = put_pack(work, pdesc, ptasks);
    lch.nptasks = int(ptasks.size());
  }

  // grouped GEMM descriptors
  std::vector<kern::GemmOut> gouts(outs.size());
  std::vector<kern::GemmPair> gpairs;
  gpairs.reserve(all_pairs.size());
  std::vector<int64_t> out_work(outs.size());
  double flops = 0, bytes = 0;
  for (size_t o = 0; o < outs.size(); ++o) {
    const Block& cb_ = out.blocks[o];
    int64_t odims[kMaxRank];
    out.shape.block_dims(cb_.idx, odims);
    int64_t M = 1, N = 1;
    for (int j = 0; j < na; ++j) M *= odims[j];
    for (int j = 0; j < nb; ++j) N *= odims[na + j];
    gouts[o].c_off = out.ptr(cb_) - ab;
    gouts[o].M = int32_t(M);
    gouts[o].N = int32_t(N);
    gouts[o].ldc = int32_t(N);
    gouts[o].pair_begin = int32_t(gpairs.size());
    int64_t ksum = 0;
    for (int q = outs[o].pb; q < outs[o].pe; ++q) {
      const auto& p = all_pairs[size_t(q)];
      const View& x = va[p.first];
      const View& y = vb[p.second];
      kern::GemmPair gp{};
      gp.a_off = x.off;
      gp.a_rs = x.rs;
      gp.a_cs = x.cs;
      gp.b_off = y.off;
      gp.b_rs = y.rs;
      gp.b_cs = y.cs;
      gp.K = int32_t(x.K);
      gpairs.push_back(gp);
      ksum += x.K;
      flops += 2.0 * double(M) * double(N) * double(x.K);
    }
    gouts[o].pair_end = int32_t(gpairs.size());
    out_work[o] = ksum;
    bytes += 8.0 * double(cb_.numel);
  }
  for (size_t k = 0; k < a.blocks.size(); ++k)
    if (used_a[k]) bytes += 8.0 * double(a.blocks[k].numel);
  for (size_t k = 0; k < b.blocks.size(); ++k)
    if (used_b[k]) bytes += 8.0 * double(b.blocks[k].numel);
  // A contraction whose blocks are all narrow (few columns, short summed K:
  // an MFMA tile would be mostly padding) runs on the row kernel (FP64
  // mode); otherwise every block takes MFMA tiles, so a contraction is one
  // launch either way.
  const char* rows_env = std::getenv("QTT_GEMM_ROWS");  // "0": MFMA tiles for every block
  const bool per_block = ctx->per_block_launch && !into && !plan;
  bool all_narrow = ctx->precision == Precision::FP64 && !per_block && !(rows_env && std::atoi(rows_env) == 0);
  for (size_t o = 0; o < outs.size() && all_narrow; ++o)
    all_narrow = gouts[o].N <= kern::kRowsMaxN && out_work[o] <= kern::kRowsMaxK;
  std::vector<char> narrow(outs.size(), char(all_narrow));
  int64_t mfma_tiles = 0;
  if (!all_narrow)
    for (size_t o = 0; o < outs.size(); ++o)
      mfma_tiles += int64_t((gouts[o].M + kern::kGemmTile - 1) / kern::kGemmTile) *
                    ((gouts[o].N + kern::kGemmTile - 1) / kern::kGemmTile);
  // Split-K when the MFMA tiles would leave most CUs idle (long-K
  // contractions with few output tiles, such as the right environment of
  // the matvec): the K extent of a block, its pairs in order, is cut into
  // about equal runs (a pair may be cut: a sub-pair is the same pair with A
  // and B advanced by k0 and a shorter K); each run accumulates into a
  // partial of its own and a fixed-order sum writes the block.  launch_outs
  // = gouts followed by the runs; split blocks get no tiles of their own.
  std::vector<kern::GemmOut> launch_outs = gouts;
  std::vector<int64_t> launch_work = out_work;
  std::vector<char> split_block(outs.size(), 0);
  std::vector<kern::SumParts> sums;
  std::vector<kern::ChunkTask> sum_tasks;
  {
    constexpr int64_t kTargetTiles = 2048, kSplitMinK = 64;
    constexpr int64_t kMaxSplit = 8;
    const char* split_env = std::getenv("QTT_GEMM_SPLIT_K");  // "0": never split
    const bool split_ok = !per_block && !(split_env && std::atoi(split_env) == 0);
    const int64_t split = split_ok ? std::min(kMaxSplit, kTargetTiles / std::max<int64_t>(mfma_tiles, 1)) : 1;
    std::vector<int> nparts(outs.size(), 0);
    int64_t part_elems = 0;
    if (split >= 2)
      for (size_t o = 0; o < outs.size(); ++o) {
        if (narrow[o]) continue;
        const int64_t np = std::min(split, out_work[o] / kSplitMinK);
        if (np < 2) continue;
        nparts[o] = int(np);
        part_elems += np * gouts[o].M * gouts[o].N;
      }
    if (part_elems) {
      int64_t pos = work.take(size_t(part_elems) * sizeof(double));
      for (size_t o = 0; o < outs.size(); ++o) {
        const int np = nparts[o];
        if (!np) continue;
        split_block[o] = 1;
        const int64_t mn = int64_t(gouts[o].M) * gouts[o].N;
        add_chunks(sum_tasks, int(sums.size()), mn);
        sums.push_back({gouts[o].c_off, pos, mn, int32_t(np), 0});
        // runs of about ksum / np, rounded up to whole k-tiles
        const int64_t run = ((out_work[o] + np - 1) / np + 15) / 16 * 16;
        int q = gouts[o].pair_begin;
        int64_t k0 = 0;
        for (int part = 0; part < np; ++part) {
          kern::GemmOut g = gouts[o];
          g.c_off = pos + int64_t(part) * mn;
          g.ldc = g.N;
          g.pair_begin = int32_t(gpairs.size());
          int64_t left = part + 1 == np ? INT64_MAX : run;
          while (left > 0 && q < gouts[o].pair_end) {
            const kern::GemmPair src = gpairs[size_t(q)];  // a copy: push_back below may reallocate
            const int64_t take = std::min<int64_t>(left, src.K - k0);
            if (take > 0) {
              kern::GemmPair sp = src;
              sp.a_off += k0 * src.a_cs;
              sp.b_off += k0 * src.b_rs;
              sp.K = int32_t(take);
              gpairs.push_back(sp);
              left -= take;
              k0 += take;
            }
            if (k0 >= src.K) {
              ++q;
              k0 = 0;
            }
          }
          g.pair_end = int32_t(gpairs.size());
          int64_t kw = 0;
          for (int r = g.pair_begin; r < g.pair_end; ++r) kw += gpairs[size_t(r)].K;
          launch_outs.push_back(g);
          launch_work.push_back(kw);
        }
        pos += int64_t(np) * mn;
      }
    }
  }
  // longest work first for a balanced tail: the tiles of one output share
  // its summed K, so ordering the outputs (stably) by it and emitting each
  // one's tiles row by row sorts the tiles by work
  std::vector<kern::GemmTile> tiles;
  std::vector<kern::RowTask> row_tasks;
  std::vector<int> ord(launch_outs.size());
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return launch_work[x] > launch_work[y]; });
  // MFMA tiles of a TM x TN variant, longest work first, row by row
  auto tiles_of = [&](int variant) {
    const int TM = kern::kVariantTM[variant], TN = kern::kVariantTN[variant];
    std::vector<kern::GemmTile> tl;
    for (int o : ord) {
      if (size_t(o) < outs.size() && (narrow[size_t(o)] || split_block[size_t(o)])) continue;
      const int tm = (launch_outs[o].M + TM - 1) / TM;
      const int tn = (launch_outs[o].N + TN - 1) / TN;
      for (int i = 0; i < tm; ++i)
        for (int j = 0; j < tn; ++j) tl.push_back({int32_t(o), int16_t(i), int16_t(j)});
    }
    return tl;
  };
  for (int o : ord)
    if (size_t(o) < outs.size() && narrow[size_t(o)]) {
      const int rpt = kern::rows_per_task(gouts[o].N);
      for (int r0 = 0; r0 < gouts[o].M; r0 += rpt) row_tasks.push_back({int32_t(o), int32_t(r0)});
    }
  // tile variant: from the tuning table for this contraction's shape class;
  // a class not seen yet is timed over every variant when the call may
  // time (not inside plans or the allocation-free path), else the 64 x 64
  // tile runs
  const bool f64_tiles = ctx->precision == Precision::FP64 && !per_block;
  int variant = kern::kTile64;
  std::string tune_key;
  bool tune_now = false;
  if (f64_tiles) {
    tune_key = gemm_shape_key(launch_outs, launch_work, mfma_tiles);
    variant = gemm_variant_for(tune_key, mfma_tiles, work.timed(), &tune_now);
  }
  tiles = tiles_of(tune_now ? int(kern::kTile64) : variant);
  if (stats) {
    stats->flops += flops;
    stats->bytes += bytes;
    stats->out_blocks += long(outs.size());
    stats->pairs += long(gpairs.size());
    stats->tiles += long(tiles.size());
  }
  if (per_block) {
    // reference-shaped baseline: one launch per block pair (mm, then addmm_);
    // the per-pair descriptors and tile lists go up in one staging copy
    std::vector<kern::GemmOut> one;
    std::vector<kern::GemmTile> tl;
    std::vector<std::pair<int, int>> launches;  // (first tile, tile count) per pair
    for (size_t o = 0; o < outs.size(); ++o) {
      const int tm = (gouts[o].M + kern::kGemmTile - 1) / kern::kGemmTile;
      const int tn = (gouts[o].N + kern::kGemmTile - 1) / kern::kGemmTile;
      if (tm * tn == 0) continue;
      for (int p = gouts[o].pair_begin; p < gouts[o].pair_end; ++p) {
        const int t0 = int(tl.size());
        for (int i = 0; i < tm; ++i)
          for (int j = 0; j < tn; ++j) tl.push_back({int32_t(one.size()), int16_t(i), int16_t(j)});
        kern::GemmOut g1 = gouts[o];
        g1.pair_begin = p;
        g1.pair_end = p + 1;
        one.push_back(g1);
        launches.push_back({t0, tm * tn});
      }
    }
    td_pack(ctx, lch, true);
    if (one.empty()) return out;
    auto [d_one, d_tl, d_pairs] = ctx->upload_pack(one, tl, gpairs);
    for (size_t l = 0; l < launches.size(); ++l) {
      const bool first = l == 0 || one[l].c_off != one[l - 1].c_off;
      KernelTimer kt(*ctx, "gemm_per_block");
      kern::gemm_f64(ctx->stream(), ab, d_tl + launches[l].first, launches[l].second, d_one, d_pairs, 1.0,
                     first ? 0.0 : 1.0);
    }
    return out;
  }
  ht_plan.stop();
  HostTimer ht_up(*ctx, work.timed() ? "host_td_upload" : nullptr);
  std::tie(lch.tiles, lch.outs, lch.pairs, lch.rows, lch.sums, lch.sum_tasks) =
      put_pack(work, tiles, launch_outs, gpairs, row_tasks, sums, sum_tasks);
  ht_up.stop();
  lch.ntiles = int(tiles.size());
  lch.nrows = int(row_tasks.size());
  lch.nsum_tasks = int(sum_tasks.size());
  lch.split_fp32 = ctx->precision == Precision::SPLIT_FP32;
  lch.variant = variant;
  lch.gemm_name = ctx->gemm_label ? ctx->gemm_label : (lch.split_fp32 ? "gemm_split" : "gemm_f64");
  if (tune_now) {
    // operands ready, then every variant timed on the real launch (each
    // writes the same output bits); the fastest is recorded and used
    td_pack(ctx, lch, false);
    lch.nptasks = 0;
    float best_ms = 0.f;
    hipEvent_t e0, e1;
    QTT_HIP(hipEventCreate(&e0));
    QTT_HIP(hipEventCreate(&e1));
    for (int v = 0; v < kern::kGemmVariants; ++v) {
      std::vector<kern::GemmTile> tl = tiles_of(v);
      const kern::GemmTile* dt = ctx->upload_vec(tl);
      float ms = 0.f;
      for (int rep = 0; rep < 2; ++rep) {
        QTT_HIP(hipEventRecord(e0, ctx->stream()));
        kern::gemm_f64_variant(ctx->stream(), v, ab, dt, int(tl.size()), lch.outs, lch.pairs, 1.0, 0.0);
        QTT_HIP(hipEventRecord(e1, ctx->stream()));
        QTT_HIP(hipEventSynchronize(e1));
        QTT_HIP(hipEventElapsedTime(&ms, e0, e1));
      }
      if (v == 0 || ms < best_ms) {
        best_ms = ms;
        lch.variant = v;
        lch.tiles = dt;
        lch.ntiles = int(tl.size());
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    gemm_variant_record(tune_key, lch.variant);
  }
  if (plan)
    *plan = lch;
  else
    td_run(ctx, lch, work.timed());
  return out;
}

}  // namespace

}  // namespace qtt
