// ttg.h -- a single-process, single-GPU implementation of the TTG API subset
// that the tiled Cholesky graph (examples/potrf of devreal/ttg) is written
// against, for MI355X.
//
// What it provides (reference interface in parentheses):
//   keys            ttg::MultiIndex<Rank>                (ttg/ttg/util/multiindex.h:12-60)
//   graph           ttg::Edge, ttg::edges, ttg::fuse, ttg::Out, ttg::make_tt,
//                   TT::in<i>/out<i>, set_keymap/priomap/devicemap/defer_writer,
//                   ttg::make_ttg, ttg::make_graph_executable, invoke()
//                                                        (make_tt.h:488-569, edge.h, tt.h:30-116)
//   sends           ttg::send<i>, ttg::sendk<i>, ttg::broadcast<i...>  (func.h:158+)
//   device tasks    ttg::device::Task (C++20 coroutine), select, wait, forward,
//                   device::send<i>, device::sendk<i>, device::broadcast<i...>,
//                   current_stream/current_device/num_devices  (device/task.h:89-696, device/device.h)
//   data            ttg::Buffer<T>, ttg::devicescratch<T>, ttg::make_scratch,
//                   ttg::scope::{Allocate, SyncIn}              (buffer.h, devicescratch.h, devicescope.h)
//   runtime         ttg::initialize/finalize/execute/fence, default_execution_context
//                                                        (run.h:30-81)
//
// How it runs (the MI355X design, not the PaRSEC one):
//   * Ready tasks sit in one priority queue ordered by the TT's priomap.
//   * Device tasks are taken from the queue in rounds: every ready device
//     task of a round runs up to its first co_await select(...), its buffers
//     are made resident in HBM (H2D on the runtime's stream when the host copy
//     is newer), then all of them are resumed inside ONE launch group of
//     libttg_chol (ttg_chol_group_begin/end): the tile kernels they call are
//     recorded and launched as one batched kernel per kind.  A round's sends
//     run when its kernels have completed (the deferred-send semantics of
//     device/task.h:205-213, 677-682).  Up to two rounds are in flight, so the
//     host prepares round r+1 while the GPU runs round r; a round that left
//     no work on the stream completes at once.
//   * Host tasks run on a pool of worker threads (ttg::initialize's thread
//     count minus the scheduling thread; run.h:30-47); a worker's sends are
//     recorded and delivered by the scheduling thread when the task ends, so
//     the graph's bookkeeping stays single-threaded.  With one thread, host
//     tasks run inline.  A buffer's host copy is brought
//     up to date when host code first asks for it (Buffer::host_ptr), so a
//     host task that only routes a tile costs no copy; the copy runs on a
//     stream of its own, after the round that last wrote the buffer (the
//     push-out of parsec/ttg.h:3524-3579, done lazily).  A host task that
//     takes a value by && and touched its host copy leaves the host copy as
//     the current one.
//   * Values travel as shared objects: a broadcast shares one value among all
//     receivers, and re-sending a task's own input forwards the same object.
//     A task that takes a value by && (a writer) while other tasks still hold
//     it waits until it is the last holder when its TT defers writers
//     (set_defer_writer(true)), and otherwise works on a copy of its own
//     (register_data_copy, parsec/ttg.h:958-1060).
//   * A device task that waits (co_await wait(...)) resumes once its round
//     has completed; its buffers are copied back on the copy stream, and the
//     kernels it launches afterwards run on a stream of their own, so neither
//     waits for the rounds issued since.
//   * ttg::cancel(info) stops the graph: no further task starts, the rounds
//     in flight complete, fence() returns, and the info is kept for the
//     caller (Runtime::error_info).
#pragma once

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cassert>
#include <condition_variable>
#include <coroutine>
#include <cstdint>
#include <deque>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <iostream>
#include <map>
#include <memory>
#include <mutex>
#include <queue>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <tuple>
#include <type_traits>
#include <unordered_map>
#include <utility>
#include <vector>

#include "ttg_chol.h"

namespace ttg {

// ---------------------------------------------------------------------------
// keys
// ---------------------------------------------------------------------------
template <std::size_t Rank, typename Int = int>
struct MultiIndex {
  static constexpr std::size_t max_index = std::size_t(1) << 21;
  std::array<Int, Rank> v{};
  MultiIndex() = default;
  template <typename... Ints, typename = std::enable_if_t<sizeof...(Ints) == Rank && (std::is_integral_v<Ints> && ...)>>
  MultiIndex(Ints... i) : v{{static_cast<Int>(i)...}} {}
  template <typename I, typename = std::enable_if_t<std::is_integral_v<I>>>
  MultiIndex(std::initializer_list<I> l) {
    std::size_t k = 0;
    for (auto x : l) v[k++] = static_cast<Int>(x);
  }
  const Int &operator[](std::size_t i) const { return v[i]; }
  Int &operator[](std::size_t i) { return v[i]; }
  // components packed in 21-bit fields, as the reference's MultiIndex::hash
  std::size_t hash() const {
    std::size_t h = 0;
    for (std::size_t i = 0; i < Rank; ++i) h = h * max_index + static_cast<std::size_t>(v[i]);
    return h;
  }
  bool operator==(const MultiIndex &o) const { return v == o.v; }
  bool operator!=(const MultiIndex &o) const { return !(v == o.v); }
  friend std::ostream &operator<<(std::ostream &os, const MultiIndex &k) {
    os << "{";
    for (std::size_t i = 0; i < Rank; ++i) os << (i ? ", " : "") << k.v[i];
    return os << "}";
  }
};

enum class ExecutionSpace { Host, CUDA, HIP, L0 };
namespace scope {
enum Scope { Allocate = 0, SyncIn = 1 };
}
using scope_t = scope::Scope;

inline bool tracing() {
  static const bool on = std::getenv("TTG_TRACE") != nullptr;
  return on;
}
template <typename... Ts>
void print(const Ts &...ts) {
  std::ostringstream os;
  (os << ... << ts);
  std::cout << os.str() << std::endl;
}

namespace detail {

// keys hash through their hash() member (MultiIndex), integral keys as themselves
template <typename K>
struct KeyHash {
  std::size_t operator()(const K &k) const {
    if constexpr (std::is_integral_v<K>)
      return static_cast<std::size_t>(k);
    else
      return k.hash();
  }
};

[[noreturn]] inline void fatal(const std::string &what) {
  std::cerr << "ttg: " << what << std::endl;
  std::abort();
}

inline void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) fatal(std::string(what) + ": " + hipGetErrorString(e));
}

// ---------------------------------------------------------------------------
// HBM pool: device blocks recycled by size
// ---------------------------------------------------------------------------
class DevicePool {
 public:
  void *get(std::size_t bytes) {
    std::lock_guard<std::mutex> g(mu_);
    auto it = free_.find(bytes);
    if (it != free_.end() && !it->second.empty()) {
      void *p = it->second.back();
      it->second.pop_back();
      return p;
    }
    void *p = nullptr;
    hip_check(hipMalloc(&p, bytes ? bytes : 1), "hipMalloc");
    ++allocated_;
    return p;
  }
  void put(void *p, std::size_t bytes) {
    std::lock_guard<std::mutex> g(mu_);
    if (p) free_[bytes].push_back(p);
  }
  void release() {
    std::lock_guard<std::mutex> g(mu_);
    for (auto &kv : free_)
      for (void *p : kv.second) (void)hipFree(p);
    free_.clear();
  }
  long allocated() const { return allocated_; }

 private:
  std::map<std::size_t, std::vector<void *>> free_;
  long allocated_ = 0;
  std::mutex mu_;  // values may die on a host worker
};

// ---------------------------------------------------------------------------
// buffers with a host copy and an HBM copy
// ---------------------------------------------------------------------------
struct BufferBase {
  void *host = nullptr;
  void *dev = nullptr;
  std::size_t bytes = 0;
  bool host_valid = true;
  bool dev_valid = false;
  uint64_t write_round = 0;  // the last scheduling round that selected the buffer for writing
  virtual ~BufferBase();
  // the host copy up to date (a push-out from HBM if the HBM copy is newer)
  void sync_host();
};

struct BufferUse {
  BufferBase *buf;
  bool readonly;
  bool sync_in;
};

}  // namespace detail

template <typename T>
class Buffer : public detail::BufferBase {
 public:
  Buffer() = default;
  explicit Buffer(std::size_t n) : owned_(new T[n]()), n_(n) { set(owned_.get(), n); }
  // non-owning view of host memory
  Buffer(T *ptr, std::size_t n) : n_(n) { set(ptr, n); }
  Buffer(const Buffer &) = delete;
  Buffer &operator=(const Buffer &) = delete;
  Buffer(Buffer &&o) noexcept { take(std::move(o)); }
  Buffer &operator=(Buffer &&o) noexcept {
    if (this != &o) {
      release_dev();
      take(std::move(o));
    }
    return *this;
  }
  ~Buffer() override { release_dev(); }

  // Host access pushes the HBM copy back first when that one is newer: a
  // host task that only routes a tile (never reads it) costs no copy.
  T *host_ptr() {
    sync_host();
    return static_cast<T *>(host);
  }
  const T *host_ptr() const {
    const_cast<Buffer *>(this)->sync_host();
    return static_cast<const T *>(host);
  }
  T *current_device_ptr() const { return static_cast<T *>(dev); }
  T *device_ptr() const { return static_cast<T *>(dev); }
  std::size_t size() const { return n_; }
  // the host contents are about to be produced on the device: skip the H2D
  // copy at the next select
  void discard_host() { host_valid = false; }
  void reset(std::size_t n) {
    release_dev();
    owned_.reset(new T[n]());
    n_ = n;
    set(owned_.get(), n);
  }

 private:
  void set(T *p, std::size_t n) {
    host = p;
    bytes = n * sizeof(T);
    host_valid = true;
    dev_valid = false;
  }
  void take(Buffer &&o) {
    owned_ = std::move(o.owned_);
    n_ = o.n_;
    host = o.host;
    dev = o.dev;
    bytes = o.bytes;
    host_valid = o.host_valid;
    dev_valid = o.dev_valid;
    o.host = o.dev = nullptr;
    o.bytes = 0;
    o.n_ = 0;
  }
  void release_dev();
  std::unique_ptr<T[]> owned_;
  std::size_t n_ = 0;
};

// Scratch memory for a device task (Allocate: no H2D; SyncIn: copied in).
template <typename T>
class devicescratch : public detail::BufferBase {
 public:
  devicescratch(T *hostp, scope_t s, std::size_t n) : scope_(s), n_(n) {
    host = hostp;
    bytes = n * sizeof(T);
    host_valid = s == scope::SyncIn;
    dev_valid = false;
  }
  devicescratch(devicescratch &&) = default;
  ~devicescratch() override;
  T *device_ptr() const { return static_cast<T *>(dev); }
  T *host_ptr() const { return static_cast<T *>(host); }
  std::size_t size() const { return n_; }
  scope_t scope() const { return scope_; }

 private:
  scope_t scope_;
  std::size_t n_;
};

template <typename T>
devicescratch<T> make_scratch(T *ptr, scope_t s, std::size_t n = 1) {
  return devicescratch<T>(ptr, s, n);
}

// Values that carry buffers expose them through ttg_buffers(value) (found by
// argument-dependent lookup); the default is none.
template <typename V>
std::vector<detail::BufferBase *> ttg_buffers(V &) {
  return {};
}

class TTBase;

namespace device {
struct Task;
using Stream = hipStream_t;
struct Device {
  int id = -1;
  bool is_device() const { return id >= 0; }
  operator int() const { return id; }
};
}  // namespace device

namespace detail {

// ---------------------------------------------------------------------------
// runtime
// ---------------------------------------------------------------------------
struct TaskBase {
  int priority = 0;
  uint64_t seq = 0;
  std::string name;
  virtual bool is_device() const = 0;
  virtual void run_host() = 0;                   // host task
  virtual std::coroutine_handle<> start() = 0;   // device task: run to first co_await
  virtual void finish() {}                       // release inputs
  // The value of a && input that other tasks still hold, when the TT defers
  // writers (set_defer_writer(true)): the task waits until it is the last
  // holder.  nullptr: nothing to wait for.
  virtual const void *writer_blocker() const { return nullptr; }
  // replace every && input that other tasks still hold by a copy of its own
  virtual void copy_shared_writers() {}
  virtual void after_host_run() {}  // host writers: their host copies are now the newer ones
  // the current task's inputs, for forwarding re-sent values without copies
  std::vector<std::pair<const void *, std::shared_ptr<void>>> inputs;
  virtual ~TaskBase() = default;
};

struct TaskOrder {
  bool operator()(const TaskBase *a, const TaskBase *b) const {
    if (a->priority != b->priority) return a->priority < b->priority;  // larger first
    return a->seq > b->seq;                                             // then FIFO
  }
};

class Runtime {
 public:
  static Runtime &get() {
    static Runtime r;
    return r;
  }
  void init(int nthreads) {
    if (inited_) return;
    nthreads_ = nthreads;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) fatal("no HIP device");
    hip_check(hipSetDevice(0), "hipSetDevice");
    hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipStreamCreateWithFlags(&copy_, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipStreamCreateWithFlags(&eager_, hipStreamNonBlocking), "hipStreamCreate");
    // TTG_ROUND_MIN / TTG_ROUND_MAX override the round sizes (below)
    if (const char *e = std::getenv("TTG_ROUND_MIN")) min_batch = std::max(1, std::atoi(e));
    if (const char *e = std::getenv("TTG_ROUND_MAX")) max_batch = std::max(1, std::atoi(e));
    // threads: the scheduling thread plus host workers; -1 takes
    // TTG_NUM_THREADS, else 5 (host tasks here are short: copies back of
    // finished tiles, routing), at most the machine's hardware threads
    int nt = nthreads;
    if (nt <= 0) {
      const char *e = std::getenv("TTG_NUM_THREADS");
      nt = e ? std::max(1, std::atoi(e)) : 5;
      nt = std::min<int>(nt, std::max(1u, std::thread::hardware_concurrency()));
    }
    nthreads_ = nt;
    worker_streams_.resize(nt - 1);
    for (auto &ws : worker_streams_)
      hip_check(hipStreamCreateWithFlags(&ws, hipStreamNonBlocking), "hipStreamCreate");
    inited_ = true;
  }
  void fini() {
    if (!inited_) return;
    fence();
    for (auto ws : worker_streams_) (void)hipStreamDestroy(ws);
    worker_streams_.clear();
    pool_.release();
    (void)hipStreamDestroy(stream_);
    (void)hipStreamDestroy(copy_);
    (void)hipStreamDestroy(eager_);
    inited_ = false;
  }
  bool inited() const { return inited_; }
  hipStream_t stream() const { return stream_; }
  DevicePool &pool() { return pool_; }
  TaskBase *current() const { return current_; }
  int num_threads() const { return nthreads_; }

  void schedule(TaskBase *t) {
    if (cancelled_) {  // a stopped graph runs nothing more
      t->finish();
      delete t;
      return;
    }
    t->seq = seq_++;
    ready_.push(t);
  }

  // Stop the graph: no task runs after the current one (ready tasks and
  // tasks that become ready are dropped, rounds in flight complete), and
  // fence() returns.  The first failure's info is kept for the caller.
  void cancel(long info) {
    if (!cancelled_) error_info_ = info;
    cancelled_ = true;
  }
  bool cancelled() const { return cancelled_; }
  long error_info() const { return error_info_; }
  // ready for the next graph
  void clear_cancel() {
    cancelled_ = false;
    error_info_ = 0;
  }

  void fence();

  // Writers (tasks taking a value by &&) when their TT defers writers: a
  // writer whose value other tasks still hold waits for it under that value
  // and is woken when a holder releases it.  Returns false when the value
  // already has a waiting writer (the caller then copies it).
  bool defer(TaskBase *t, const void *value) {
    if (waiting_.count(value)) return false;
    waiting_.emplace(value, t);
    return true;
  }
  bool has_waiting() const { return !waiting_.empty(); }
  // a holder of `value` let it go: its waiting writer, if any, runs now or
  // waits for the next value it shares
  void released(const void *value) {
    auto it = waiting_.find(value);
    if (it == waiting_.end()) return;
    TaskBase *t = it->second;
    waiting_.erase(it);
    admit(t);
  }
  // a task taken from the ready queue: false if it had to wait (deferred)
  bool admit_check(TaskBase *t) {
    const void *v = t->writer_blocker();
    if (!v) return true;
    if (!defer(t, v)) {
      t->copy_shared_writers();
      return true;
    }
    return false;
  }
  void admit(TaskBase *t) {
    if (cancelled_) {
      t->finish();
      delete t;
      return;
    }
    if (admit_check(t)) ready_.push(t);
  }

  // statistics
  long tasks_run = 0, rounds = 0, kernel_launches = 0;
  std::atomic<long> push_outs{0};
  long max_round = 0;
  int max_inflight = 2;
  // Round sizes: forming a round costs host time per task (~1 us) while the
  // GPU may idle, so a round that starts an idle GPU takes at most min_batch
  // ready tasks, and while earlier rounds run each round may take 1.5 times
  // the last one (the host forms round r + 1 as the GPU runs round r), up to
  // max_batch.
  int min_batch = 512;
  int max_batch = 1 << 20;
  int last_batch_ = 0;

  // buffer validity flags: the scheduling thread and host workers
  // (push-outs) both change them
  std::mutex &flag_mutex() { return flag_mu_; }

  void make_resident(BufferBase *b, bool sync_in) {
    if (!b->dev) b->dev = pool_.get(b->bytes);
    if (sync_in && b->host_valid && !b->dev_valid && b->bytes)
      hip_check(hipMemcpyAsync(b->dev, b->host, b->bytes, hipMemcpyHostToDevice, stream_), "H2D");
    if (sync_in) b->dev_valid = true;
  }
  // device -> host, synchronous, on a copy stream of the calling thread;
  // ordered after the round in flight that writes the buffer, if any (not
  // behind every round queued since)
  void push_out(BufferBase *b) {
    std::unique_lock<std::mutex> lk(flag_mu_);
    if (!(b->dev && b->dev_valid && !b->host_valid && b->host && b->bytes)) {
      b->host_valid = true;
      return;
    }
    const uint64_t wr = b->write_round;
    lk.unlock();
    hipStream_t cs = worker_copy_ ? worker_copy_ : copy_;
    if (wr > completed_round_.load()) {
      if (!worker_copy_) {
        bool found = false;
        for (auto &r : inflight_)
          if (r.id == wr) {
            found = true;
            if (r.done) hip_check(hipStreamWaitEvent(cs, r.done, 0), "hipStreamWaitEvent");
          }
        // the round being formed writes it: its kernels are not even queued
        if (!found) fatal("host access to a buffer that a device task of the round being formed writes");
      } else {
        // a worker waits for the scheduling thread to complete that round
        std::unique_lock<std::mutex> rl(round_mu_);
        round_cv_.wait(rl, [&] { return completed_round_.load() >= wr; });
      }
    }
    hip_check(hipMemcpyAsync(b->host, b->dev, b->bytes, hipMemcpyDeviceToHost, cs), "D2H");
    hip_check(hipStreamSynchronize(cs), "hipStreamSynchronize");
    lk.lock();
    b->host_valid = true;
    ++push_outs;
  }
  void release(BufferBase *b) {
    if (b->dev) pool_.put(b->dev, b->bytes);
    b->dev = nullptr;
    b->dev_valid = false;
  }

  // the task whose body runs on this thread
  static inline thread_local TaskBase *current_ = nullptr;
  // sends of a host task on a worker, delivered by the scheduling thread
  static inline thread_local std::vector<std::function<void()>> *captured_ = nullptr;

 private:
  static inline thread_local hipStream_t worker_copy_ = nullptr;
  // The workers live for one fence(): started when it begins and joined
  // before it returns, so no thread outlives the graph (or the process's
  // own teardown).
  void start_workers() {
    if (!workers_.empty() || worker_streams_.empty()) return;
    stop_ = false;
    for (hipStream_t ws : worker_streams_) workers_.emplace_back([this, ws] { worker_main(ws); });
  }
  void stop_workers() {
    if (workers_.empty()) return;
    {
      std::lock_guard<std::mutex> g(pool_mu_);
      stop_ = true;
    }
    pool_cv_.notify_all();
    for (auto &w : workers_) w.join();
    workers_.clear();
  }
  void worker_main(hipStream_t ws) {
    hip_check(hipSetDevice(0), "hipSetDevice");
    worker_copy_ = ws;
    for (;;) {
      TaskBase *t = nullptr;
      {
        std::unique_lock<std::mutex> lk(pool_mu_);
        pool_cv_.wait(lk, [&] { return stop_ || !host_q_.empty(); });
        if (host_q_.empty()) break;  // stopping
        t = host_q_.front();
        host_q_.pop_front();
      }
      std::vector<std::function<void()>> sends;
      captured_ = &sends;
      current_ = t;
      t->run_host();
      t->after_host_run();
      current_ = nullptr;
      captured_ = nullptr;
      {
        std::lock_guard<std::mutex> g(pool_mu_);
        done_q_.emplace_back(t, std::move(sends));
      }
      done_cv_.notify_one();
    }
    worker_copy_ = nullptr;
  }
  // host tasks the workers finished: deliver their sends, release inputs
  void collect(bool block) {
    std::vector<std::pair<TaskBase *, std::vector<std::function<void()>>>> done;
    {
      std::unique_lock<std::mutex> lk(pool_mu_);
      if (block) done_cv_.wait(lk, [&] { return !done_q_.empty(); });
      done.swap(done_q_);
    }
    for (auto &[t, sends] : done) {
      for (auto &f : sends) f();
      t->finish();
      delete t;
      ++tasks_run;
      --outstanding_;
    }
  }
  void run_host_task(TaskBase *t) {
    if (workers_.empty()) {
      current_ = t;
      t->run_host();
      t->after_host_run();
      current_ = nullptr;
      t->finish();
      delete t;
      ++tasks_run;
      return;
    }
    ++outstanding_;
    {
      std::lock_guard<std::mutex> g(pool_mu_);
      host_q_.push_back(t);
    }
    pool_cv_.notify_one();
  }

  struct Round {
    std::vector<std::pair<TaskBase *, std::coroutine_handle<>>> tasks;
    hipEvent_t done = nullptr;
    uint64_t id = 0;
  };
  void complete(Round &r);
  void drop_ready();
  void run_device_round(std::vector<TaskBase *> &batch);

  bool inited_ = false;
  int nthreads_ = 1;
  hipStream_t stream_ = nullptr;
  // push-outs to host tasks: their data comes from rounds that have
  // completed, so the copy need not wait behind the rounds queued since
  hipStream_t copy_ = nullptr;
  // kernels a device task launches after a wait: its data is complete, so
  // they need not queue behind the rounds issued since
  hipStream_t eager_ = nullptr;
  DevicePool pool_;
  // deferred writers, each under the value it waits for
  std::unordered_map<const void *, TaskBase *> waiting_;
  uint64_t round_seq_ = 0;
  std::atomic<uint64_t> completed_round_{0};
  std::mutex flag_mu_, round_mu_;
  std::condition_variable round_cv_;
  // host workers
  std::vector<std::thread> workers_;
  std::vector<hipStream_t> worker_streams_;  // one copy stream per worker
  std::mutex pool_mu_;
  std::condition_variable pool_cv_, done_cv_;
  std::deque<TaskBase *> host_q_;
  std::vector<std::pair<TaskBase *, std::vector<std::function<void()>>>> done_q_;
  long outstanding_ = 0;  // host tasks handed to workers, not yet collected
  bool stop_ = false;
  bool cancelled_ = false;
  long error_info_ = 0;
  std::priority_queue<TaskBase *, std::vector<TaskBase *>, TaskOrder> ready_;
  std::vector<Round> inflight_;
  uint64_t seq_ = 0;
  std::vector<hipEvent_t> events_;
};

inline BufferBase::~BufferBase() = default;

inline void BufferBase::sync_host() {
  if (!host_valid && dev && dev_valid && Runtime::get().inited()) Runtime::get().push_out(this);
}

}  // namespace detail

template <typename T>
void Buffer<T>::release_dev() {
  if (dev && detail::Runtime::get().inited()) detail::Runtime::get().release(this);
  dev = nullptr;
}

template <typename T>
devicescratch<T>::~devicescratch() {
  if (dev && detail::Runtime::get().inited()) detail::Runtime::get().release(this);
}

// ---------------------------------------------------------------------------
// device-task coroutine protocol
// ---------------------------------------------------------------------------
namespace device {

namespace detail {
inline thread_local Stream current_stream_v = nullptr;
inline thread_local int current_device_v = -1;
}  // namespace detail

inline Stream current_stream() { return detail::current_stream_v; }
inline Device current_device() { return Device{detail::current_device_v}; }
inline int num_devices() {
  int n = 0;
  return hipGetDeviceCount(&n) == hipSuccess ? n : 0;
}

enum class State { Init, Select, Wait, SendOut, Done };

struct select_t {
  std::vector<ttg::detail::BufferUse> uses;
};
struct wait_t {
  std::vector<ttg::detail::BufferBase *> bufs;
};
struct send_t {
  std::function<void()> action;
};
struct forward_t {
  std::vector<std::function<void()>> actions;
};

namespace detail {
template <typename B>
ttg::detail::BufferUse use_of(B &b) {
  using D = std::remove_reference_t<B>;
  constexpr bool ro = std::is_const_v<D>;
  auto *base = const_cast<ttg::detail::BufferBase *>(static_cast<const ttg::detail::BufferBase *>(&b));
  return ttg::detail::BufferUse{base, ro, true};
}
template <typename T>
ttg::detail::BufferUse use_of(ttg::devicescratch<T> &s) {
  return ttg::detail::BufferUse{&s, false, s.scope() == ttg::scope::SyncIn};
}
}  // namespace detail

// Make the listed Buffers / scratch resident on the device and select the
// stream; resumes inside the launch group of the current round.
template <typename... Args>
[[nodiscard]] select_t select(Args &&...args) {
  select_t s;
  (s.uses.push_back(detail::use_of(args)), ...);
  return s;
}

// Wait for the kernels of this task and copy the listed buffers to the host.
template <typename... Args>
[[nodiscard]] wait_t wait(Args &&...args) {
  wait_t w;
  (w.bufs.push_back(const_cast<ttg::detail::BufferBase *>(static_cast<const ttg::detail::BufferBase *>(&args))),
   ...);
  return w;
}

template <typename... Sends>
[[nodiscard]] forward_t forward(Sends &&...s) {
  forward_t f;
  (f.actions.push_back(std::move(s.action)), ...);
  return f;
}

struct Task {
  struct promise_type {
    State state = State::Init;
    select_t sel;
    wait_t wt;
    std::vector<std::function<void()>> sends;

    Task get_return_object() { return Task{std::coroutine_handle<promise_type>::from_promise(*this)}; }
    std::suspend_never initial_suspend() noexcept { return {}; }
    std::suspend_always final_suspend() noexcept {
      state = State::Done;
      return {};
    }
    void return_void() {}
    void unhandled_exception() { ttg::detail::fatal("exception escaped a device task"); }

    std::suspend_always await_transform(select_t &&s) {
      sel = std::move(s);
      state = State::Select;
      return {};
    }
    std::suspend_always await_transform(wait_t &&w) {
      wt = std::move(w);
      state = State::Wait;
      return {};
    }
    std::suspend_always await_transform(forward_t &&f) {
      for (auto &a : f.actions) sends.push_back(std::move(a));
      state = State::SendOut;
      return {};
    }
    std::suspend_always await_transform(send_t &&s) {
      sends.push_back(std::move(s.action));
      state = State::SendOut;
      return {};
    }
  };
  std::coroutine_handle<promise_type> h;
};

}  // namespace device

// ---------------------------------------------------------------------------
// edges and terminals
// ---------------------------------------------------------------------------
namespace detail {

template <typename Key, typename Value>
struct InTerminal {
  std::function<void(const Key &, std::shared_ptr<Value>)> deliver;
};
template <typename Key>
struct InTerminal<Key, void> {
  std::function<void(const Key &)> deliver;
};

template <typename Key, typename Value>
struct EdgeImpl {
  std::string name;
  std::vector<InTerminal<Key, Value> *> consumers;
};

}  // namespace detail

template <typename Key, typename Value = void>
class Edge {
 public:
  using key_type = Key;
  using value_type = Value;
  explicit Edge(const std::string &name = "") : impl_(std::make_shared<detail::EdgeImpl<Key, Value>>()) {
    impl_->name = name;
  }
  std::shared_ptr<detail::EdgeImpl<Key, Value>> impl() const { return impl_; }

 private:
  std::shared_ptr<detail::EdgeImpl<Key, Value>> impl_;
};

// several edges feeding one input terminal
template <typename Key, typename Value>
struct FusedEdges {
  std::vector<std::shared_ptr<detail::EdgeImpl<Key, Value>>> impls;
};

template <typename Key, typename Value, typename... More>
FusedEdges<Key, Value> fuse(const Edge<Key, Value> &e, const More &...more) {
  FusedEdges<Key, Value> f;
  f.impls.push_back(e.impl());
  (f.impls.push_back(more.impl()), ...);
  return f;
}

template <typename... E>
auto edges(E &&...e) {
  return std::make_tuple(std::forward<E>(e)...);
}

template <typename Key, typename Value = void>
class Out {
 public:
  using key_type = Key;
  using value_type = Value;
  Out() = default;
  void connect(std::shared_ptr<detail::EdgeImpl<Key, Value>> e) { edges_.push_back(std::move(e)); }

  // a host task on a worker thread records its sends; the scheduling
  // thread delivers them when the task ends
  template <typename V = Value, typename = std::enable_if_t<!std::is_void_v<V>>>
  void send_shared(const Key &k, const std::shared_ptr<V> &v) {
    if (auto *cap = detail::Runtime::captured_) {
      cap->push_back([this, k, v] { this->send_shared(k, v); });
      return;
    }
    for (auto &e : edges_)
      for (auto *c : e->consumers) c->deliver(k, v);
  }
  void sendk(const Key &k) {
    static_assert(std::is_void_v<Value>, "sendk needs a void-valued terminal");
    if (auto *cap = detail::Runtime::captured_) {
      cap->push_back([this, k] { this->sendk(k); });
      return;
    }
    for (auto &e : edges_)
      for (auto *c : e->consumers) c->deliver(k);
  }

 private:
  std::vector<std::shared_ptr<detail::EdgeImpl<Key, Value>>> edges_;
};

namespace detail {

// Share a value being sent: an input of the running task is forwarded as is
// (no copy), an rvalue is moved into a new shared object, an lvalue copied.
template <typename Value, typename V>
std::shared_ptr<Value> share(V &&v) {
  using D = std::decay_t<V>;
  static_assert(std::is_same_v<D, Value>, "sent value type must match the terminal");
  if (TaskBase *t = Runtime::get().current())
    for (auto &in : t->inputs)
      if (in.first == static_cast<const void *>(&v)) return std::static_pointer_cast<Value>(in.second);
  if constexpr (std::is_rvalue_reference_v<V &&>)
    return std::make_shared<Value>(std::move(v));
  else
    return std::make_shared<Value>(v);
}

}  // namespace detail

template <std::size_t i, typename Key, typename V, typename... Outs>
void send(const Key &k, V &&v, std::tuple<Outs...> &out) {
  using O = std::tuple_element_t<i, std::tuple<Outs...>>;
  auto sp = detail::share<typename O::value_type>(std::forward<V>(v));
  std::get<i>(out).send_shared(k, sp);
}

template <std::size_t i, typename Key, typename... Outs>
void sendk(const Key &k, std::tuple<Outs...> &out) {
  std::get<i>(out).sendk(k);
}

namespace detail {
template <typename KL>
auto as_keylist(const KL &kl) {
  if constexpr (std::is_class_v<KL> && requires(const KL &x) { x.begin(); }) {
    return std::vector<typename KL::value_type>(kl.begin(), kl.end());
  } else {
    return std::vector<KL>{kl};
  }
}
template <std::size_t N, std::size_t I, std::size_t... Is, typename Tup, typename Value, typename... Outs>
void broadcast_impl(const Tup &keylists, const std::shared_ptr<Value> &v, std::tuple<Outs...> &out) {
  for (const auto &k : as_keylist(std::get<N>(keylists))) std::get<I>(out).send_shared(k, v);
  if constexpr (sizeof...(Is) > 0) broadcast_impl<N + 1, Is...>(keylists, v, out);
}
}  // namespace detail

// broadcast<i, j, ...>(tuple(keylist_i, keylist_j, ...), value, out): one
// shared value to every key of every listed terminal.
template <std::size_t I, std::size_t... Is, typename Tup, typename V, typename... Outs>
void broadcast(const Tup &keylists, V &&v, std::tuple<Outs...> &out) {
  using O = std::tuple_element_t<I, std::tuple<Outs...>>;
  auto sp = detail::share<typename O::value_type>(std::forward<V>(v));
  detail::broadcast_impl<0, I, Is...>(keylists, sp, out);
}

namespace device {
// Deferred forms: executed after the task's kernels complete.
template <std::size_t i, typename Key, typename V, typename... Outs>
[[nodiscard]] send_t send(const Key &k, V &&v, std::tuple<Outs...> &out) {
  using O = std::tuple_element_t<i, std::tuple<Outs...>>;
  auto sp = ttg::detail::share<typename O::value_type>(std::forward<V>(v));
  return send_t{[k, sp, &out]() { std::get<i>(out).send_shared(k, sp); }};
}
template <std::size_t i, typename Key, typename... Outs>
[[nodiscard]] send_t sendk(const Key &k, std::tuple<Outs...> &out) {
  return send_t{[k, &out]() { std::get<i>(out).sendk(k); }};
}
template <std::size_t I, std::size_t... Is, typename Tup, typename V, typename... Outs>
[[nodiscard]] send_t broadcast(Tup keylists, V &&v, std::tuple<Outs...> &out) {
  using O = std::tuple_element_t<I, std::tuple<Outs...>>;
  auto sp = ttg::detail::share<typename O::value_type>(std::forward<V>(v));
  return send_t{[keylists = std::move(keylists), sp, &out]() {
    ttg::detail::broadcast_impl<0, I, Is...>(keylists, sp, out);
  }};
}
}  // namespace device

// ---------------------------------------------------------------------------
// template tasks
// ---------------------------------------------------------------------------
class TTBase {
 public:
  explicit TTBase(std::string name) : name_(std::move(name)) {}
  virtual ~TTBase() = default;
  const std::string &get_name() const { return name_; }
  virtual long tasks_created() const { return created_; }

 protected:
  std::string name_;
  long created_ = 0;
};

namespace detail {

template <typename F>
struct fn_traits : fn_traits<decltype(&F::operator())> {};
template <typename C, typename R, typename... A>
struct fn_traits<R (C::*)(A...) const> {
  using ret = R;
  using args = std::tuple<A...>;
};
template <typename C, typename R, typename... A>
struct fn_traits<R (C::*)(A...)> {
  using ret = R;
  using args = std::tuple<A...>;
};

template <typename Tup>
struct last_of;
template <typename... A>
struct last_of<std::tuple<A...>> {
  using type = std::tuple_element_t<sizeof...(A) - 1, std::tuple<A...>>;
};

template <typename Tup, typename Seq>
struct slice;
template <typename Tup, std::size_t... I>
struct slice<Tup, std::index_sequence<I...>> {
  using type = std::tuple<std::tuple_element_t<I, Tup>...>;
};

template <typename Tup, std::size_t B, std::size_t E>
using slice_t = typename slice<Tup, decltype([]<std::size_t... I>(std::index_sequence<I...>) {
  return std::index_sequence<(B + I)...>{};
}(std::make_index_sequence<E - B>{}))>::type;

// registration of an input terminal on every edge feeding it
template <typename Key, typename Value>
void attach(const Edge<Key, Value> &e, InTerminal<Key, Value> *t) {
  e.impl()->consumers.push_back(t);
}
template <typename Key, typename Value>
void attach(const FusedEdges<Key, Value> &f, InTerminal<Key, Value> *t) {
  for (auto &i : f.impls) i->consumers.push_back(t);
}
template <typename Key, typename Value>
void attach_out(const Edge<Key, Value> &e, Out<Key, Value> &o) {
  o.connect(e.impl());
}

}  // namespace detail

// TT with a non-void key and one or more value inputs.
template <typename Key, typename OutTuple, typename Func, typename... In>
class TT : public TTBase {
  static constexpr std::size_t NIn = sizeof...(In);
  using values_t = std::tuple<std::decay_t<In>...>;
  using ret_t = std::invoke_result_t<Func, const Key &, In..., OutTuple &>;
  static constexpr bool is_device = std::is_same_v<ret_t, device::Task>;

  struct Inst : detail::TaskBase {
    TT *tt = nullptr;
    Key key{};
    std::array<std::shared_ptr<void>, NIn> in{};
    std::size_t pending = NIn;
    bool is_device() const override { return TT::is_device; }

    template <std::size_t... I>
    decltype(auto) call(std::index_sequence<I...>) {
      return tt->f_(key, arg<I>()..., tt->outs_);
    }
    template <std::size_t I>
    decltype(auto) arg() {
      using A = std::tuple_element_t<I, std::tuple<In...>>;
      using V = std::tuple_element_t<I, values_t>;
      V &v = *std::static_pointer_cast<V>(in[I]);
      if constexpr (std::is_rvalue_reference_v<A>)
        return std::move(v);
      else
        return static_cast<A>(v);
    }
    template <std::size_t I>
    static constexpr bool writes() {
      return std::is_rvalue_reference_v<std::tuple_element_t<I, std::tuple<In...>>>;
    }
    void bind_inputs() {
      if (!tt->defer_writer_) copy_shared_writers();
      inputs.clear();
      for (auto &p : in) inputs.emplace_back(p.get(), p);
    }
    const void *writer_blocker() const override {
      if (!tt->defer_writer_) return nullptr;
      return blocker(std::make_index_sequence<NIn>{});
    }
    template <std::size_t... I>
    const void *blocker(std::index_sequence<I...>) const {
      const void *v = nullptr;
      ((v = v ? v : (writes<I>() && in[I].use_count() > 1 ? in[I].get() : nullptr)), ...);
      return v;
    }
    void copy_shared_writers() override { copy_writers(std::make_index_sequence<NIn>{}); }
    template <std::size_t... I>
    void copy_writers(std::index_sequence<I...>) {
      (
          [&] {
            if constexpr (writes<I>()) {
              using V = std::tuple_element_t<I, values_t>;
              if (in[I].use_count() > 1) {
                if constexpr (std::is_copy_constructible_v<V>)
                  in[I] = std::make_shared<V>(*std::static_pointer_cast<V>(in[I]));
                else
                  detail::fatal(tt->name_ + ": a writer of a shared value that cannot be copied");
              }
            }
          }(),
          ...);
    }
    void run_host() override {
      bind_inputs();
      if constexpr (!TT::is_device) call(std::make_index_sequence<NIn>{});
    }
    // a host writer that touched a buffer's host copy (host_ptr) leaves that
    // copy as the current one
    void after_host_run() override { invalidate_dev(std::make_index_sequence<NIn>{}); }
    template <std::size_t... I>
    void invalidate_dev(std::index_sequence<I...>) {
      (
          [&] {
            if constexpr (writes<I>()) {
              using V = std::tuple_element_t<I, values_t>;
              if (in[I]) {
                std::lock_guard<std::mutex> g(detail::Runtime::get().flag_mutex());
                for (auto *b : ttg_buffers(*std::static_pointer_cast<V>(in[I])))
                  if (b->host_valid) b->dev_valid = false;
              }
            }
          }(),
          ...);
    }
    std::coroutine_handle<> start() override {
      bind_inputs();
      if constexpr (TT::is_device) {
        device::Task t = call(std::make_index_sequence<NIn>{});
        return t.h;
      } else {
        return {};
      }
    }
    // release the inputs; a value now held by one other task may let a
    // writer waiting for it run
    void finish() override {
      inputs.clear();
      auto &rt = detail::Runtime::get();
      for (auto &p : in) {
        if (p && p.use_count() == 2 && rt.has_waiting()) {
          const void *v = p.get();
          p.reset();
          rt.released(v);
        } else {
          p.reset();
        }
      }
    }
  };

 public:
  TT(Func f, std::string name, int) : TTBase(std::move(name)), f_(std::move(f)) {}

  template <typename InEdges, typename OutEdges>
  void wire(const InEdges &ins, const OutEdges &outs) {
    wire_in(ins, std::make_index_sequence<NIn>{});
    wire_out(outs, std::make_index_sequence<std::tuple_size_v<OutTuple>>{});
  }

  template <std::size_t i>
  auto *in() {
    return &std::get<i>(termstore_);
  }
  template <std::size_t i>
  auto &out() {
    return std::get<i>(outs_);
  }

  template <typename F>
  void set_keymap(F &&) {}  // one process: every key is local
  template <typename F>
  void set_devicemap(F &&) {}  // one GPU
  void set_defer_writer(bool b) { defer_writer_ = b; }
  template <typename F>
  void set_priomap(F &&f) {
    prio_ = std::forward<F>(f);
  }

  long tasks_created() const override { return created_; }

 private:
  template <std::size_t I, typename V>
  void deliver(const Key &k, std::shared_ptr<V> v) {
    auto &slot = pending_[k];
    if (!slot) {
      slot.reset(new Inst);
      slot->tt = this;
      slot->key = k;
      slot->name = name_;
    }
    if (slot->in[I]) detail::fatal(name_ + ": input delivered twice for one key");
    slot->in[I] = std::move(v);
    if (--slot->pending == 0) {
      Inst *t = slot.release();
      pending_.erase(k);
      t->priority = prio_ ? prio_(k) : 0;
      ++created_;
      detail::Runtime::get().schedule(t);
    }
  }

  template <typename InEdges, std::size_t... I>
  void wire_in(const InEdges &ins, std::index_sequence<I...>) {
    (
        [&] {
          using V = std::tuple_element_t<I, values_t>;
          auto &t = std::get<I>(termstore_);
          t.deliver = [this](const Key &k, std::shared_ptr<V> v) { this->deliver<I, V>(k, std::move(v)); };
          detail::attach(std::get<I>(ins), &t);
        }(),
        ...);
  }
  template <typename OutEdges, std::size_t... I>
  void wire_out(const OutEdges &outs, std::index_sequence<I...>) {
    (detail::attach_out(std::get<I>(outs), std::get<I>(outs_)), ...);
  }

  Func f_;
  OutTuple outs_;
  std::tuple<detail::InTerminal<Key, std::decay_t<In>>...> termstore_;
  std::unordered_map<Key, std::unique_ptr<Inst>, detail::KeyHash<Key>> pending_;
  std::function<int(const Key &)> prio_;
  bool defer_writer_ = false;  // the reference's default (TTG_PARSEC_DEFER_WRITER, parsec/ttg.h:13)
};

// TT without a key and without inputs (started by invoke()).
template <typename OutTuple, typename Func>
class TTVoid : public TTBase {
 public:
  TTVoid(Func f, std::string name) : TTBase(std::move(name)), f_(std::move(f)) {}
  template <typename OutEdges>
  void wire(const OutEdges &outs) {
    wire_out(outs, std::make_index_sequence<std::tuple_size_v<OutTuple>>{});
  }
  void invoke() {
    ++created_;
    f_(outs_);
  }
  template <std::size_t i>
  auto &out() {
    return std::get<i>(outs_);
  }
  template <typename F>
  void set_keymap(F &&) {}
  void set_defer_writer(bool) {}

 private:
  template <typename OutEdges, std::size_t... I>
  void wire_out(const OutEdges &outs, std::index_sequence<I...>) {
    (detail::attach_out(std::get<I>(outs), std::get<I>(outs_)), ...);
  }
  Func f_;
  OutTuple outs_;
};

// TT whose inputs are void-valued (sendk) edges: func(const Key&, outs&);
// a host or a device task.
template <typename Key, typename OutTuple, typename Func>
class TTKeyOnly : public TTBase {
  using ret_t = std::invoke_result_t<Func, const Key &, OutTuple &>;
  static constexpr bool is_dev = std::is_same_v<ret_t, device::Task>;

 public:
  TTKeyOnly(Func f, std::string name) : TTBase(std::move(name)), f_(std::move(f)) {}
  template <typename InEdges, typename OutEdges>
  void wire(const InEdges &ins, const OutEdges &outs) {
    term_.deliver = [this](const Key &k) {
      ++created_;
      auto *t = new Inst;
      t->tt = this;
      t->key = k;
      t->name = name_;
      t->priority = prio_ ? prio_(k) : 0;
      detail::Runtime::get().schedule(t);
    };
    detail::attach(std::get<0>(ins), &term_);
    wire_out(outs, std::make_index_sequence<std::tuple_size_v<OutTuple>>{});
  }
  template <std::size_t i>
  auto &out() {
    return std::get<i>(outs_);
  }
  template <typename F>
  void set_keymap(F &&) {}
  void set_defer_writer(bool) {}
  template <typename F>
  void set_priomap(F &&f) {
    prio_ = std::forward<F>(f);
  }

 private:
  struct Inst : detail::TaskBase {
    TTKeyOnly *tt = nullptr;
    Key key{};
    bool is_device() const override { return is_dev; }
    void run_host() override {
      if constexpr (!is_dev) tt->f_(key, tt->outs_);
    }
    std::coroutine_handle<> start() override {
      if constexpr (is_dev) {
        device::Task t = tt->f_(key, tt->outs_);
        return t.h;
      } else {
        return {};
      }
    }
  };
  template <typename OutEdges, std::size_t... I>
  void wire_out(const OutEdges &outs, std::index_sequence<I...>) {
    (detail::attach_out(std::get<I>(outs), std::get<I>(outs_)), ...);
  }
  Func f_;
  OutTuple outs_;
  detail::InTerminal<Key, void> term_;
  std::function<int(const Key &)> prio_;
};

namespace detail {
template <typename T>
struct is_out_tuple : std::false_type {};
template <typename... O>
struct is_out_tuple<std::tuple<O...>> : std::true_type {};
}  // namespace detail

// make_tt<ES>(func, edges(inputs...), edges(outputs...), name, innames, outnames)
template <ExecutionSpace ES = ExecutionSpace::Host, typename Func, typename InEdges, typename OutEdges>
auto make_tt(Func f, const InEdges &ins, const OutEdges &outs, const std::string &name = "TT",
             const std::vector<std::string> & = {}, const std::vector<std::string> & = {}) {
  using args = typename detail::fn_traits<Func>::args;
  constexpr std::size_t NA = std::tuple_size_v<args>;
  using out_ref = std::tuple_element_t<NA - 1, args>;
  using OutTuple = std::remove_cvref_t<out_ref>;
  if constexpr (NA == 1) {
    auto tt = std::make_unique<TTVoid<OutTuple, Func>>(std::move(f), name);
    tt->wire(outs);
    return tt;
  } else {
    using KeyArg = std::tuple_element_t<0, args>;
    using Key = std::remove_cvref_t<KeyArg>;
    if constexpr (NA == 2) {
      auto tt = std::make_unique<TTKeyOnly<Key, OutTuple, Func>>(std::move(f), name);
      tt->wire(ins, outs);
      return tt;
    } else {
      using Ins = detail::slice_t<args, 1, NA - 1>;
      return [&]<typename... In>(std::tuple<In...> *) {
        auto tt = std::make_unique<TT<Key, OutTuple, Func, In...>>(std::move(f), name, 0);
        tt->wire(ins, outs);
        return tt;
      }(static_cast<Ins *>(nullptr));
    }
  }
}

// make_tt<KeyT>(...): the key type spelled explicitly (the reference's
// ttg::make_tt<void>(...) for keyless start-up tasks); runs on the host.
template <typename KeyT, typename Func, typename InEdges, typename OutEdges>
auto make_tt(Func f, const InEdges &ins, const OutEdges &outs, const std::string &name = "TT",
             const std::vector<std::string> &inn = {}, const std::vector<std::string> &outn = {}) {
  return make_tt<ExecutionSpace::Host>(std::move(f), ins, outs, name, inn, outn);
}

// a composite of TTs (the reference's ttg::TTG, tt.h:30-116)
class TTG : public TTBase {
 public:
  TTG(std::vector<std::unique_ptr<TTBase>> ops, std::string name) : TTBase(std::move(name)), ops_(std::move(ops)) {}
  const std::vector<std::unique_ptr<TTBase>> &ops() const { return ops_; }

 private:
  std::vector<std::unique_ptr<TTBase>> ops_;
};

template <typename Ins, typename Outs>
auto make_ttg(std::vector<std::unique_ptr<TTBase>> &&ops, const Ins &, const Outs &, const std::string &name = "TTG") {
  return std::make_unique<TTG>(std::move(ops), name);
}

inline bool make_graph_executable(TTBase *) { return true; }

// ---------------------------------------------------------------------------
// world
// ---------------------------------------------------------------------------
struct World {
  int rank() const { return 0; }
  int size() const { return 1; }
  void profile_on() {}
  void profile_off() {}
  void dag_on(const char *) {}
  void dag_off() {}
};
inline World default_execution_context() { return World{}; }
inline World ttg_default_execution_context() { return World{}; }

inline void initialize(int, char **, int num_threads = -1) { detail::Runtime::get().init(num_threads); }
inline void finalize() { detail::Runtime::get().fini(); }
inline void execute(World = {}) {}
inline void fence(World = {}) { detail::Runtime::get().fence(); }
// ends the process, as the reference's ttg::abort (run.h:62)
[[noreturn]] inline void abort() { std::abort(); }
// Stop the running graph cleanly with `info` (for instance LAPACK's info of
// a failed factorization): no further task starts, fence() returns, and
// ttg::cancelled() / ttg::cancel_info() report it until the next
// ttg::clear_cancel().
inline void cancel(long info) { detail::Runtime::get().cancel(info); }
inline bool cancelled() { return detail::Runtime::get().cancelled(); }
inline long cancel_info() { return detail::Runtime::get().error_info(); }
inline void clear_cancel() { detail::Runtime::get().clear_cancel(); }

// ---------------------------------------------------------------------------
// the scheduler loop
// ---------------------------------------------------------------------------
namespace detail {

inline void Runtime::run_device_round(std::vector<TaskBase *> &batch) {
  using P = device::Task::promise_type;
  Round r;
  r.id = ++round_seq_;
  device::detail::current_stream_v = stream_;
  device::detail::current_device_v = 0;
  // phase 1: every task runs to its select; buffers made resident
  std::vector<std::coroutine_handle<P>> hs;
  for (TaskBase *t : batch) {
    current_ = t;
    auto h = std::coroutine_handle<P>::from_address(t->start().address());
    current_ = nullptr;
    P &p = h.promise();
    if (p.state == device::State::Select)
      for (auto &u : p.sel.uses) {
        std::lock_guard<std::mutex> g(flag_mu_);
        make_resident(u.buf, u.sync_in);
        // the HBM copy is the current one from here on: the task's kernels
        // write every buffer it did not select read-only
        if (!u.readonly) {
          u.buf->dev_valid = true;
          u.buf->host_valid = false;
          u.buf->write_round = r.id;
        }
      }
    else if (p.state != device::State::Done)
      fatal(t->name + ": a device task must co_await select(...) before anything else");
    hs.push_back(h);
  }
  // phase 2: resume everything inside one launch group
  if (ttg_chol_group_begin(stream_) != TTG_CHOL_OK) fatal(ttg_chol_last_error());
  for (std::size_t i = 0; i < batch.size(); ++i) {
    auto h = hs[i];
    P &p = h.promise();
    if (p.state == device::State::Select) {
      current_ = batch[i];
      h.resume();
      current_ = nullptr;
      if (p.state == device::State::Select) fatal(batch[i]->name + ": a second select in one device task");
    }
    r.tasks.emplace_back(batch[i], h);
  }
  if (ttg_chol_group_end(stream_) != TTG_CHOL_OK) fatal(ttg_chol_last_error());
  kernel_launches += ttg_chol_group_last_launches();
  // a round that left no work on the stream (nothing launched or copied, and
  // everything before it done) is complete now: no event, no wait
  const hipError_t q = hipStreamQuery(stream_);
  if (q != hipSuccess && q != hipErrorNotReady) hip_check(q, "hipStreamQuery");
  if (q == hipErrorNotReady) {
    if (events_.empty()) {
      hipEvent_t e;
      hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
      events_.push_back(e);
    }
    r.done = events_.back();
    events_.pop_back();
    hip_check(hipEventRecord(r.done, stream_), "hipEventRecord");
  }
  ++rounds;
  max_round = std::max<long>(max_round, (long)batch.size());
  inflight_.push_back(std::move(r));
}

inline void Runtime::complete(Round &r) {
  using P = device::Task::promise_type;
  if (r.done) {
    hip_check(hipEventSynchronize(r.done), "hipEventSynchronize");
    events_.push_back(r.done);
    r.done = nullptr;
  }
  {
    std::lock_guard<std::mutex> g(round_mu_);
    completed_round_ = r.id;
  }
  if (!workers_.empty()) round_cv_.notify_all();
  for (auto &[t, hv] : r.tasks) {
    auto h = std::coroutine_handle<P>::from_address(hv.address());
    P &p = h.promise();
    // drive the task to its end: waits copy back and resume; sends (deferred
    // until the task's kernels completed, which they have) run, then resume
    while (p.state != device::State::Done) {
      const bool after_wait = p.state == device::State::Wait;
      if (after_wait) {
        // the task's kernels are done (its round is): copy back on the copy
        // stream, without waiting for the rounds issued since
        bool copied = false;
        for (auto *b : p.wt.bufs) {
          if (b->dev && b->host && b->bytes) {
            hip_check(hipMemcpyAsync(b->host, b->dev, b->bytes, hipMemcpyDeviceToHost, copy_), "D2H");
            copied = true;
          }
          b->host_valid = true;
        }
        if (copied) hip_check(hipStreamSynchronize(copy_), "hipStreamSynchronize");
      } else if (p.state == device::State::SendOut) {
        current_ = t;
        for (auto &s : p.sends) s();
        current_ = nullptr;
        p.sends.clear();
      } else {
        fatal(t->name + ": device task suspended outside select/wait/send");
      }
      p.state = device::State::Init;
      current_ = t;
      // kernels issued after a wait run eagerly (not grouped) on a stream of
      // their own and complete before the task's next sends
      if (after_wait) device::detail::current_stream_v = eager_;
      h.resume();
      device::detail::current_stream_v = stream_;
      current_ = nullptr;
      if (after_wait) hip_check(hipStreamSynchronize(eager_), "hipStreamSynchronize");
    }
    h.destroy();
    t->finish();
    delete t;
    ++tasks_run;
  }
}

inline void Runtime::drop_ready() {
  while (!ready_.empty()) {
    TaskBase *t = ready_.top();
    ready_.pop();
    t->finish();
    delete t;
  }
  std::vector<TaskBase *> w;
  for (auto &kv : waiting_) w.push_back(kv.second);
  waiting_.clear();
  for (TaskBase *t : w) {
    t->finish();
    delete t;
  }
}

inline void Runtime::fence() {
  start_workers();
  for (;;) {
    if (outstanding_) collect(false);
    if (cancelled_) drop_ready();
    // host tasks first (they feed the device rounds), highest priority first
    std::vector<TaskBase *> dev_batch;
    const int limit = inflight_.empty() ? min_batch
                                        : std::min(max_batch, std::max(min_batch, last_batch_ + last_batch_ / 2));
    while (!ready_.empty() && !cancelled_) {
      TaskBase *t = ready_.top();
      if (t->is_device() && (int)dev_batch.size() >= limit) break;
      ready_.pop();
      if (!admit_check(t)) continue;  // a writer waiting for its readers
      if (t->is_device()) {
        dev_batch.push_back(t);
        continue;
      }
      run_host_task(t);
    }
    if (cancelled_) {
      for (TaskBase *t : dev_batch) {
        t->finish();
        delete t;
      }
      dev_batch.clear();
    }
    if (!dev_batch.empty()) {
      if ((int)inflight_.size() >= max_inflight) {
        complete(inflight_.front());
        inflight_.erase(inflight_.begin());
      }
      last_batch_ = (int)dev_batch.size();
      run_device_round(dev_batch);
      continue;
    }
    if (inflight_.empty()) {
      if (outstanding_) {  // host tasks still running on workers
        collect(true);
        continue;
      }
      if (waiting_.empty() || cancelled_) break;
      // writers still waiting with nothing left to run: the holders of their
      // values never run, so the writers take copies and go
      std::vector<TaskBase *> w;
      for (auto &kv : waiting_) w.push_back(kv.second);
      waiting_.clear();
      for (TaskBase *t : w) {
        t->copy_shared_writers();
        ready_.push(t);
      }
      continue;
    }
    complete(inflight_.front());
    inflight_.erase(inflight_.begin());
  }
  if (cancelled_) drop_ready();
  stop_workers();
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
}

}  // namespace detail
}  // namespace ttg
