// Device runtime: the HBM arena that owns all block data of one context, the
// context (device, HIP stream, arena, descriptor staging) and timing hooks.
#pragma once

#include <chrono>
#include <map>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <cstring>
#include <tuple>
#include <vector>

#include "qtt_common.hpp"

namespace qtt {

// One hipMalloc'd region suballocated first-fit with 256-byte alignment and
// coalescing of free neighbours.  Allocation is per tensor (all blocks of a
// tensor are contiguous), so the free list stays short.
class Arena {
 public:
  explicit Arena(size_t bytes);
  ~Arena();
  Arena(const Arena&) = delete;
  Arena& operator=(const Arena&) = delete;

  size_t alloc(size_t bytes);  // byte offset from base(); throws OutOfMemory
  void release(size_t off);
  char* base() const { return base_; }
  size_t capacity() const { return cap_; }
  size_t used() const { return used_; }
  size_t peak() const { return peak_; }
  void reset_peak() { peak_ = used_; }

 private:
  char* base_ = nullptr;
  size_t cap_ = 0, used_ = 0, peak_ = 0;
  std::map<size_t, size_t> free_;            // offset -> size
  std::unordered_map<size_t, size_t> live_;  // offset -> size
  std::mutex mu_;
};

// A region of an arena.  It shares ownership of the arena, so a tensor that
// outlives its context still releases its bytes into live memory (the arena
// is freed with its last region).
struct DeviceBuffer {
  std::shared_ptr<Arena> arena;
  size_t off;
  size_t bytes;
  DeviceBuffer(std::shared_ptr<Arena> a, size_t nbytes)
      : arena(std::move(a)), off(nbytes ? arena->alloc(nbytes) : 0), bytes(nbytes) {}
  ~DeviceBuffer() {
    if (bytes) arena->release(off);
  }
  double* dptr() const { return reinterpret_cast<double*>(arena->base() + off); }
  char* cptr() const { return arena->base() + off; }
};
using BufPtr = std::shared_ptr<DeviceBuffer>;

enum class Precision : int {
  FP64 = 0,        // FP64 MFMA contractions
  SPLIT_FP32 = 1,  // hi/lo split FP32 MFMA, compensated, recombined in FP64
};

struct KernelTimer;

class Context {
 public:
  Context(int device, size_t arena_bytes, size_t stage_bytes = size_t(64) << 20);
  ~Context();
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;

  int device() const { return device_; }
  hipStream_t stream() const { return stream_; }
  // A second stream for independent work beside the main one, created on
  // the first fork() (a context that never forks keeps one stream, and so
  // one hardware queue): fork() makes it wait for everything enqueued on the
  // main stream so far, join() makes the main stream wait for everything
  // enqueued on it.  Both are device-side waits (no host synchronization).
  hipStream_t side_stream() const { return side_; }
  void fork();
  void join();
  Arena& arena() { return *arena_; }
  BufPtr alloc(size_t bytes) { return std::make_shared<DeviceBuffer>(arena_, bytes); }

  // Copies host descriptors to device memory in stream order; returns a device
  // pointer valid for kernels enqueued afterwards on this context's stream.
  void* upload(const void* data, size_t bytes);
  template <class T>
  T* upload_vec(const std::vector<T>& v) {
    return static_cast<T*>(upload(v.data(), v.size() * sizeof(T)));
  }
  // Several descriptor lists in one staging copy (one host-to-device
  // transfer instead of one per list); each list starts 16-byte aligned.
  template <class... T>
  std::tuple<T*...> upload_pack(const std::vector<T>&... v) {
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
    char* d = static_cast<char*>(upload(buf.data(), total));
    size_t i = 0;
    return std::tuple<T*...>{reinterpret_cast<T*>(d + offs[i++])...};
  }
  // Copies host bytes through the pinned staging ring to any device address
  // (stream-ordered; no allocation).
  void stage_to(void* dev_dst, const void* data, size_t bytes);
  // This device's copy of the static scratch of the allocation-free
  // contraction (its address is taken once per device when the first
  // context on it is created).
  void* noalloc_scratch() const { return noalloc_scratch_; }
  void sync();
  // Waits for an event recorded on this stream (a blocking wait: the host
  // thread sleeps instead of polling, which matters when several solve
  // threads wait at once).
  void wait_event(hipEvent_t e);

  // Pinned host slots for small device-to-host reads that the host waits on
  // with read_event(slot) while later work keeps the stream busy: slot 0 and
  // slot 1 each hold n_per_slot 32-bit words (grown on demand; growing
  // synchronizes the stream).
  // Device-to-host copy of a few bytes through pinned memory, waited for
  // (the stream is drained up to the copy).  A copy into pageable memory
  // would go through the runtime's own staging.
  void read_small(void* host_dst, const void* dev_src, size_t bytes);
  // A pinned word that a device-to-host copy may target without waiting;
  // valid after the next synchronization of the stream.
  int32_t* pinned_flag() { return host_small_; }
  void reserve_read(size_t n_per_slot);
  int32_t* read_slot(int slot) const { return host_read_ + size_t(slot) * read_cap_; }
  hipEvent_t read_event(int slot) const { return read_ev_[slot]; }

  // Handle object of the C ABI that owns this context (opaque here).
  void* owner = nullptr;
  // Block SVD health over the context's life: calls, the largest number of
  // sweeps/Gram iterations one call needed, and sectors whose Jacobi sweeps
  // reached the sweep cap still rotating.
  int64_t svd_calls = 0, svd_max_sweeps = 0, svd_unconverged = 0;
  // SVD model FLOPs (sum over sectors of 14 m n^2 + 8 n^3, m >= n) of every
  // block SVD, and of the sectors whose rotations ran as multi-work-group
  // rounds
  double svd_model_flops = 0, svd_model_flops_wide = 0;
  Precision precision = Precision::FP64;
  // Two-site eigensolver of DMRG: 0 = the reference's one-step Lanczos;
  // k >= 2 = k-step Lanczos with full reorthogonalisation, residual tolerance
  // lanczos_tol (relative to the eigenvalue).
  int lanczos_krylov = 0;
  double lanczos_tol = 1e-12;
  int64_t lanczos_updates = 0, lanczos_vectors = 0;  // k-step mode: updates and Krylov vectors built
  // k-step mode: the matvec of an update is captured in a HIP graph and
  // replayed for the later Krylov vectors (count of replays below)
  bool graphs = true;
  int64_t graph_replays = 0;
  // Reference-shaped baseline: one GEMM launch per block pair (mm, addmm_).
  bool per_block_launch = false;
  // Independent work of one call may run on the side stream.  Processes that
  // run several contexts at once turn it off: their streams already share the
  // device, and a side stream takes a hardware queue that another context's
  // stream then waits behind.
  bool overlap = true;
  // Per-kernel accumulated time is recorded when profiling is on.
  bool profiling = false;
  // Timer name of the contraction GEMMs (callers label their call sites;
  // null: "gemm_f64" / "gemm_split").
  const char* gemm_label = nullptr;
  std::map<std::string, std::pair<double, long>> kernel_ms;  // name -> (ms, launches)
  void record(const char* name, hipEvent_t a, hipEvent_t b);
  hipEvent_t event();
  void flush_timings();  // synchronizes and folds pending event pairs into kernel_ms

 private:
  int device_;
  hipStream_t stream_ = nullptr;
  hipStream_t side_ = nullptr;
  hipEvent_t fork_ev_ = nullptr, join_ev_ = nullptr;
  std::shared_ptr<Arena> arena_;
  char* host_stage_ = nullptr;
  BufPtr dev_stage_;
  size_t stage_cap_ = 0, stage_pos_ = 0;
  std::vector<hipEvent_t> event_pool_;
  size_t event_next_ = 0;
  struct Pending {
    const char* name;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending_;
  int32_t* host_small_ = nullptr;  // word 0: pinned_flag; from byte 256: read_small
  static constexpr size_t kSmallBytes = size_t(64) << 10;
  int32_t* host_read_ = nullptr;
  size_t read_cap_ = 0;  // words per slot
  hipEvent_t read_ev_[2] = {nullptr, nullptr};
  hipEvent_t sync_ev_ = nullptr;
  void* noalloc_scratch_ = nullptr;
  char* stage_host(const void* data, size_t bytes, size_t* pos);
};

// RAII helper: brackets one launch with events when profiling is enabled.
struct KernelTimer {
  Context& ctx;
  const char* name;
  hipEvent_t a = nullptr, b = nullptr;
  KernelTimer(Context& c, const char* n) : ctx(c), name(n) {
    if (ctx.profiling) {
      a = ctx.event();
      b = ctx.event();
      QTT_HIP(hipEventRecord(a, ctx.stream()));
    }
  }
  ~KernelTimer() {
    if (ctx.profiling) {
      (void)hipEventRecord(b, ctx.stream());
      ctx.record(name, a, b);
    }
  }
};

// RAII helper: wall time of a host-side phase (sector matching, descriptor
// building) when profiling is enabled, folded into kernel_ms under a name
// starting with "host_".  stop() ends the phase early.
struct HostTimer {
  Context& ctx;
  const char* name;
  bool on;
  std::chrono::steady_clock::time_point t0;
  HostTimer(Context& c, const char* n) : ctx(c), name(n), on(c.profiling && n != nullptr) {
    if (on) t0 = std::chrono::steady_clock::now();
  }
  void stop() {
    if (!on) return;
    on = false;
    auto& e = ctx.kernel_ms[name];
    e.first += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    e.second += 1;
  }
  ~HostTimer() { stop(); }
};

}  // namespace qtt
