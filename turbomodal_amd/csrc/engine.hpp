// Internal structures of the MI355X cyclic modal engine (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <condition_variable>
#include <cstdint>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/tmhip.h"

namespace tme {

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct ArgError : std::invalid_argument {
  using std::invalid_argument::invalid_argument;
};
struct NumericError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define TM_HIP(call)                                                              \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess)                                                         \
      throw ::tme::HipError(std::string(#call) + ": " + hipGetErrorString(e_));    \
  } while (0)

// Cache of device blocks owned by one context.  Every user of a context's
// blocks launches on the context's single stream, so a block released by one
// batch and handed to the next is only touched by work ordered after every
// earlier use: no synchronisation is needed on reuse, and the steady state of
// a long sweep makes no hipMalloc / hipFree calls.
class DevPool {
 public:
  DevPool() = default;
  DevPool(const DevPool&) = delete;
  DevPool& operator=(const DevPool&) = delete;
  ~DevPool() { trim(); }
  static size_t round(size_t n) {
    const size_t g = n <= (size_t(1) << 20) ? 4096 : (size_t(2) << 20);
    return (n + g - 1) / g * g;
  }
  // a cached block of at least n bytes (n is rounded up), or nullptr
  void* take(size_t& n) {
    n = round(n);
    auto it = free_.lower_bound(n);
    if (it == free_.end() || it->first > 2 * n + (size_t(4) << 20)) return nullptr;
    void* p = it->second;
    n = it->first;
    cached_ -= n;
    free_.erase(it);
    return p;
  }
  void give(void* p, size_t n) {
    free_.emplace(n, p);
    cached_ += n;
  }
  void trim() {
    for (auto& f : free_) (void)hipFree(f.second);
    free_.clear();
    cached_ = 0;
  }
  size_t cached() const { return cached_; }

 private:
  std::multimap<size_t, void*> free_;
  size_t cached_ = 0;
};

// Device buffer owned by the engine (from a context's pool when one is given).
// `bytes` is the size asked for (what copies may move); `cap` the block held.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0, cap = 0;
  DevPool* pool = nullptr;
  DevBuf() = default;
  explicit DevBuf(size_t n) { alloc(n); }
  DevBuf(DevPool* pl, size_t n) { alloc(pl, n); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes), cap(o.cap), pool(o.pool) {
    o.p = nullptr;
    o.bytes = o.cap = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p; bytes = o.bytes; cap = o.cap; pool = o.pool;
      o.p = nullptr; o.bytes = o.cap = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t n) { alloc(nullptr, n); }
  void alloc(DevPool* pl, size_t n) {
    release();
    pool = pl;
    if (n == 0) return;
    size_t m = n;
    if (pool) {
      if ((p = pool->take(m))) {  // m: the cached block's size
        bytes = n;
        cap = m;
        return;
      }
    }
    hipError_t e = hipMalloc(&p, m);
    if (e != hipSuccess && pool && pool->cached()) {
      (void)hipGetLastError();
      pool->trim();
      e = hipMalloc(&p, m);
    }
    if (e != hipSuccess) {
      p = nullptr;
      throw HipError("hipMalloc(" + std::to_string(m) + " B): " + hipGetErrorString(e));
    }
    bytes = n;
    cap = m;
  }
  void release() {
    if (p) {
      if (pool) pool->give(p, cap);
      else (void)hipFree(p);
    }
    p = nullptr;
    bytes = cap = 0;
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

template <class T>
inline DevBuf upload(DevPool* pool, const std::vector<T>& v, hipStream_t s) {
  DevBuf b(pool, v.size() * sizeof(T));
  if (!v.empty()) {
    TM_HIP(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
    TM_HIP(hipStreamSynchronize(s));  // the host vector may be a temporary
  }
  return b;
}

template <class T>
inline DevBuf upload(const std::vector<T>& v, hipStream_t s) {
  return upload<T>(nullptr, v, s);
}

// Device-to-host delivery of results off the compute stream.  A kernel writes a
// staging block on the compute stream; a worker thread waits for that work on
// its own stream, copies the block into the caller's host buffer and recycles
// it, while the compute stream goes on with the next batch.  Callbacks queued
// with then() run on the worker once everything queued before them is done.
class Deliverer {
 public:
  explicit Deliverer(int device);
  ~Deliverer();
  Deliverer(const Deliverer&) = delete;
  Deliverer& operator=(const Deliverer&) = delete;
  // a staging block of at least n bytes (cap: its size); waits while more than
  // kMaxInFlight bytes are waiting for their copies
  void* stage(size_t n, size_t& cap);
  void unstage(void* p, size_t cap);  // a block that will not be copied
  // copy n bytes of src to host dst once the work queued so far on s is done
  void copy(hipStream_t s, void* src, size_t cap, void* dst, size_t n);
  // fn("") after the work queued so far on s and every earlier copy; fn(message) on a copy error
  void then(hipStream_t s, std::function<void(const std::string&)> fn);
  void drain();        // waits for every queued job; throws the first copy error
  void drain_quiet();  // waits, keeps the error for the next drain
  static constexpr size_t kMaxInFlight = size_t(1) << 30;

 private:
  struct Job {
    hipEvent_t ev = nullptr;
    void* src = nullptr;
    size_t cap = 0;
    void* dst = nullptr;
    size_t n = 0;
    std::function<void(const std::string&)> fn;
  };
  void enqueue(hipStream_t s, Job j);
  void run();
  int device_;
  hipStream_t cs_ = nullptr;  // the worker's copies
  hipStream_t ws_ = nullptr;  // waits for the compute streams' points (queued by their threads)
  hipEvent_t mark_ = nullptr;
  std::thread th_;
  std::mutex mu_;
  std::condition_variable cv_, cv_done_;
  std::deque<Job> q_;
  std::multimap<size_t, void*> free_;
  std::vector<hipEvent_t> events_;
  size_t inflight_ = 0;
  int active_ = 0;
  bool stop_ = false;
  std::string err_;
};

// Per-kernel event timing (bench roofline) -----------------------------------
struct KernelStat {
  int64_t launches = 0;
  double total_ms = 0, bytes = 0, flops = 0;
};

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  // device address of the library's static scratch array on this context's device
  // (hipGetSymbolAddress at context creation), for the allocation-free entry point
  void* scratch = nullptr;
  size_t scratch_bytes = 0;
  DevPool pool;  // workspaces of the solve batches
  std::string last_error;
  bool profiling = false;
  // share of the free device memory (plus this context's cached blocks) that a
  // solve call sizes its batches against: 1 - memory reserve fraction
  double batch_fraction = 0.85;
  std::map<std::string, KernelStat> stats;
  // pending timed launches: (name, start, stop, bytes, flops)
  struct Pending { std::string name; hipEvent_t a, b; double bytes, flops; };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> event_pool;
  std::unique_ptr<Deliverer> deliver;  // created on first use
  // contexts of the slot groups that one solve call runs side by side (own
  // stream, pool and deliverer), kept so their pools' cached workspaces serve
  // the next call; their deleters synchronise and release them
  std::vector<std::shared_ptr<Context>> groups;
  // captured launch sequences (hipGraph executables) by the bytes of every argument
  // of their launches (run_captured); a replay issues exactly what a fresh recording
  // would.  Released with the context.
  std::map<std::string, hipGraphExec_t> graphs;
  void clear_graphs() {
    for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
    graphs.clear();
  }
  hipEvent_t get_event();
  void flush_timing();  // synchronises and folds pending events into stats
  // A second stream (default priority) and timing-free events for launch sequences that
  // fork from `stream` and join it again within one recorded sequence; both created on
  // first use and released with the context (release_side).
  hipStream_t side = nullptr;
  std::vector<hipEvent_t> fork_events;
  // the factorization forks its shared fronts and inversions onto the side stream (set
  // while one slot group has the device: beside other groups' streams the hardware
  // queues are shared and the fork measured no faster)
  bool fork = false;
  hipStream_t side_stream();
  hipEvent_t fork_event(size_t i);
  void release_side();
  Deliverer& deliverer() {
    if (!deliver) deliver = std::make_unique<Deliverer>(device);
    return *deliver;
  }
};

// Runs record() (kernel launches on ctx->stream only: no synchronisation, copies,
// memsets or allocations) through ctx's graph cache: the first call with a key
// captures the launches into a hipGraph and instantiates it, later calls with the
// same key replay it.  `key` must hold every value the launches depend on.
// TMHIP_GRAPHS=0 in the environment launches directly instead.
void run_captured(Context* ctx, const std::string& key, const std::function<void()>& record);

// Appends the bytes of v to a graph key.
template <class T>
inline void key_add(std::string& k, const T& v) {
  k.append(reinterpret_cast<const char*>(&v), sizeof(T));
}

// RAII scope that adds its host wall time to stats["host:" + name] (profiling only).
struct HostTimer {
  Context* ctx;
  const char* name;
  double t0 = 0;
  HostTimer(Context* c, const char* n);
  ~HostTimer();
};

// RAII scope that times everything launched on ctx->stream inside it.
struct KTimer {
  Context* ctx;
  std::string name;
  double bytes, flops;
  hipEvent_t a = nullptr;
  KTimer(Context* c, const char* n, double by = 0, double fl = 0);
  ~KTimer();
};

// Symbolic analysis of one mesh ----------------------------------------------
struct Snode {
  int col0 = 0;        // first permuted free node of the pivot block
  int ncol = 0;        // pivot nodes
  int nstr = 0;        // nodes below the pivot block
  int parent = -1;
  int height = 0;
  int child[2] = {-1, -1};
  int64_t foff = 0;    // offset (complex entries) of the m x m front
  int64_t toff = 0;    // offset (complex entries) of the r-long update vector
  int str0 = 0;        // offset into struct_nodes
  int rel0 = 0;        // offset into relmap (node-level positions in parent front)
  int asm0 = 0, asm1 = 0;  // range in asm lists
  int m() const { return 3 * (ncol + nstr); }
  int w() const { return 3 * ncol; }
  int r() const { return 3 * nstr; }
};

struct Symbolic {
  int nn = 0, ne = 0, N = 0, axis = 2;
  int c1 = 0, c2 = 1, cax = 2;
  std::vector<int> elements;          // ne*10
  std::vector<double> coords;         // nn*3
  std::vector<int> node_class, partner;
  std::vector<uint8_t> cmask;         // 3*nn
  // full node -> permuted free node (interior/left nodes and right nodes via partner), -1 removed
  std::vector<int> free_of_full;
  std::vector<int> free_to_full;      // nf: interior/left full node id
  std::vector<uint8_t> pinned;        // 3*nf
  int nf = 0;
  int n_free_dofs = 0;                // reference n_free
  int n_reduced_nodes = 0;
  // node-block CSR pattern of K(k) over permuted free nodes
  std::vector<int> rowptr, colidx, tslot;
  // segmented assembly: per slot, contributions eb*4 + flags  (eb = e*100 + a*10 + b)
  std::vector<int> kc_ptr, kc_src, kp_ptr, kp_src;
  // centrifugal load: per free node, contributions (e*10+a)*2 + rotated
  std::vector<int> f_ptr, f_src;
  // fronts
  std::vector<Snode> sn;
  std::vector<int> struct_nodes;      // concatenated struct lists (permuted free nodes)
  std::vector<int> relmap;            // per snode struct node -> node position in parent front
  std::vector<int> asm_slot, asm_pos; // per snode: (slot of (I,J)), (rowpos<<16 | colpos) node-level
  std::vector<std::vector<int>> levels;   // snodes by height
  int64_t front_entries = 0;          // sum m^2
  int64_t tvec_entries = 0;           // sum r
  int64_t nnz_L = 0;
  double flops = 0;
  int max_front = 0;
  // reference free-DOF ordering (interior DOFs then left DOFs, constrained removed):
  // for each such DOF, its system index 3*perm_node + c
  std::vector<int> ref_free_to_sys;
};

void build_symbolic(Symbolic& S);
// Per front: shared (1) when neither it nor any front below it assembles an entry of a
// slot with phase contributions (its own or its transposed slot): A(k) = K + s dK - sigma M
// is then real on the whole subtree and the same for every harmonic, and so are the
// front's factor and update matrix.
std::vector<char> shared_fronts(const Symbolic& S);

}  // namespace tme
