#include "runtime.hpp"

#include "kernels.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>

namespace qtt {

std::string charge_str(const Groups& g, const Charge& c) {
  std::string s = "(";
  for (int i = 0; i < g.n; ++i) {
    if (i) s += ",";
    s += std::to_string(c.v[i]);
  }
  return s + ")";
}

static constexpr size_t kAlign = 256;
static size_t round_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

Arena::Arena(size_t bytes) {
  cap_ = round_up(bytes);
  QTT_HIP(hipMalloc(&base_, cap_));
  free_[0] = cap_;
}

Arena::~Arena() {
  if (base_) (void)hipFree(base_);
}

size_t Arena::alloc(size_t bytes) {
  size_t need = round_up(bytes ? bytes : 1);
  std::lock_guard<std::mutex> lk(mu_);
  for (auto it = free_.begin(); it != free_.end(); ++it) {
    if (it->second >= need) {
      size_t off = it->first;
      size_t rest = it->second - need;
      free_.erase(it);
      if (rest) free_[off + need] = rest;
      live_[off] = need;
      used_ += need;
      if (used_ > peak_) peak_ = used_;
      return off;
    }
  }
  throw OutOfMemory("device arena exhausted: requested " + std::to_string(need) + " bytes, " +
                    std::to_string(cap_ - used_) + " free of " + std::to_string(cap_));
}

void Arena::release(size_t off) {
  std::lock_guard<std::mutex> lk(mu_);
  auto lv = live_.find(off);
  if (lv == live_.end()) return;
  size_t sz = lv->second;
  live_.erase(lv);
  used_ -= sz;
  auto nx = free_.lower_bound(off);
  // merge with the next free range
  if (nx != free_.end() && nx->first == off + sz) {
    sz += nx->second;
    nx = free_.erase(nx);
  }
  // merge with the previous free range
  if (nx != free_.begin()) {
    auto pv = std::prev(nx);
    if (pv->first + pv->second == off) {
      pv->second += sz;
      return;
    }
  }
  free_[off] = sz;
}

Context::Context(int device, size_t arena_bytes, size_t stage_bytes) : device_(device) {
  QTT_HIP(hipSetDevice(device));
  // library set-up on this device: the address of its static scratch
  QTT_HIP(kern::noalloc_scratch_address(device, &noalloc_scratch_));
  QTT_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  QTT_HIP(hipEventCreateWithFlags(&fork_ev_, hipEventDisableTiming));
  QTT_HIP(hipEventCreateWithFlags(&join_ev_, hipEventDisableTiming));
  arena_ = std::make_shared<Arena>(arena_bytes);
  // the descriptor staging ring comes out of the arena: at most a quarter of it
  stage_cap_ = round_up(std::min(stage_bytes, std::max<size_t>(size_t(1) << 20, arena_bytes / 4)));
  QTT_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_stage_), stage_cap_, hipHostMallocDefault));
  dev_stage_ = alloc(stage_cap_);
  for (auto& e : read_ev_) QTT_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  QTT_HIP(hipEventCreateWithFlags(&sync_ev_, hipEventDisableTiming));
  QTT_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_small_), kSmallBytes, hipHostMallocDefault));
}

Context::~Context() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  if (side_) (void)hipStreamSynchronize(side_);
  pending_.clear();
  if (fork_ev_) (void)hipEventDestroy(fork_ev_);
  if (join_ev_) (void)hipEventDestroy(join_ev_);
  for (auto e : event_pool_) (void)hipEventDestroy(e);
  for (auto e : read_ev_)
    if (e) (void)hipEventDestroy(e);
  if (sync_ev_) (void)hipEventDestroy(sync_ev_);
  if (host_read_) (void)hipHostFree(host_read_);
  if (host_small_) (void)hipHostFree(host_small_);
  dev_stage_.reset();
  arena_.reset();
  if (host_stage_) (void)hipHostFree(host_stage_);
  if (side_) (void)hipStreamDestroy(side_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

// Reserves `bytes` of the staging ring at *pos and fills its host half.
char* Context::stage_host(const void* data, size_t bytes, size_t* pos) {
  size_t need = round_up(bytes ? bytes : 1);
  if (need > stage_cap_) throw OutOfMemory("descriptor upload larger than the staging ring");
  if (stage_pos_ + need > stage_cap_) {
    // wrap: everything previously staged must have been consumed, by the
    // work of both streams
    QTT_HIP(hipStreamSynchronize(stream_));
    if (side_) QTT_HIP(hipStreamSynchronize(side_));
    stage_pos_ = 0;
  }
  *pos = stage_pos_;
  char* h = host_stage_ + stage_pos_;
  if (bytes) std::memcpy(h, data, bytes);
  stage_pos_ += need;
  return h;
}

void* Context::upload(const void* data, size_t bytes) {
  size_t pos = 0;
  char* h = stage_host(data, bytes, &pos);
  char* d = dev_stage_->cptr() + pos;
  if (bytes) QTT_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, stream_));
  return d;
}

void Context::stage_to(void* dev_dst, const void* data, size_t bytes) {
  size_t pos = 0;
  char* h = stage_host(data, bytes, &pos);
  if (bytes) QTT_HIP(hipMemcpyAsync(dev_dst, h, bytes, hipMemcpyHostToDevice, stream_));
}

void Context::fork() {
  if (!side_) QTT_HIP(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking));
  QTT_HIP(hipEventRecord(fork_ev_, stream_));
  QTT_HIP(hipStreamWaitEvent(side_, fork_ev_, 0));
}

void Context::join() {
  QTT_HIP(hipEventRecord(join_ev_, side_));
  QTT_HIP(hipStreamWaitEvent(stream_, join_ev_, 0));
}

void Context::sync() {
  QTT_HIP(hipEventRecord(sync_ev_, stream_));
  wait_event(sync_ev_);
}

void Context::wait_event(hipEvent_t e) { QTT_HIP(hipEventSynchronize(e)); }

void Context::read_small(void* host_dst, const void* dev_src, size_t bytes) {
  constexpr size_t kOff = 256;
  char* dst = static_cast<char*>(host_dst);
  const char* src = static_cast<const char*>(dev_src);
  while (bytes) {
    const size_t n = std::min(bytes, kSmallBytes - kOff);
    char* h = reinterpret_cast<char*>(host_small_) + kOff;
    QTT_HIP(hipMemcpyAsync(h, src, n, hipMemcpyDeviceToHost, stream_));
    sync();
    std::memcpy(dst, h, n);
    dst += n;
    src += n;
    bytes -= n;
  }
}

void Context::reserve_read(size_t n_per_slot) {
  if (n_per_slot > read_cap_) {
    QTT_HIP(hipStreamSynchronize(stream_));  // no read may still target the old area
    if (host_read_) QTT_HIP(hipHostFree(host_read_));
    host_read_ = nullptr;
    const size_t cap = std::max<size_t>(n_per_slot, 1024);
    QTT_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_read_), 2 * cap * sizeof(int32_t), hipHostMallocDefault));
    read_cap_ = cap;
  }
}

hipEvent_t Context::event() {
  if (event_next_ == event_pool_.size()) {
    hipEvent_t e;
    QTT_HIP(hipEventCreate(&e));
    event_pool_.push_back(e);
  }
  return event_pool_[event_next_++];
}

void Context::record(const char* name, hipEvent_t a, hipEvent_t b) {
  pending_.push_back({name, a, b});
  if (pending_.size() > 4096) flush_timings();
}

void Context::flush_timings() {
  QTT_HIP(hipStreamSynchronize(stream_));
  for (auto& p : pending_) {
    float ms = 0.f;
    QTT_HIP(hipEventElapsedTime(&ms, p.a, p.b));
    auto& e = kernel_ms[p.name];
    e.first += ms;
    e.second += 1;
  }
  pending_.clear();
  event_next_ = 0;
}

}  // namespace qtt
