// Runtime of libvulkpy_hip: error state, device contexts, the stream-ordered
// caching allocator, events (jobs), host/device copies, fills and graphs.
//
// Replaces the Vulkan objects of vulkpy/_vkarray.cc: GPU (460-574),
// Buffer<T> (38-130) and Job (392-457).  Where the reference creates a
// pipeline, descriptor pool, command pool and fence per dispatch and host-waits
// every dependency (_vkarray.cc:403-438), everything here goes to one in-order
// HIP stream per context and a job is a single event.
#include "common.h"

#include <chrono>
#include <thread>

namespace vkh {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }

// ---- allocator -------------------------------------------------------------
size_t Allocator::round_size(size_t n) {
  if (n == 0) n = 1;
  const size_t small = 512, large = size_t(2) << 20;
  if (n < (size_t(1) << 20)) return ceil_div(n, small) * small;
  return ceil_div(n, large) * large;
}

void* Allocator::allocate(size_t nbytes) {
  size_t sz = round_size(nbytes);
  std::lock_guard<std::mutex> lk(mu_);
  auto it = free_.find(sz);
  if (it != free_.end() && !it->second.empty()) {
    void* p = it->second.back();
    it->second.pop_back();
    cached_bytes_ -= sz;
    live_[p] = sz;
    live_bytes_ += sz;
    return p;
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, sz);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    // Give every cached block back and try once more.  hipFree waits for the
    // device, so no enqueued kernel still uses the released blocks.
    for (auto& kv : free_)
      for (void* q : kv.second) (void)hipFree(q);
    free_.clear();
    cached_bytes_ = 0;
    e = hipMalloc(&p, sz);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      fail(VKH_OUT_OF_MEMORY, "device out of memory allocating " + std::to_string(sz) + " bytes");
    }
  }
  live_[p] = sz;
  live_bytes_ += sz;
  return p;
}

void Allocator::release(void* p) {
  if (!p) return;
  std::lock_guard<std::mutex> lk(mu_);
  auto it = live_.find(p);
  if (it == live_.end()) fail(VKH_INVALID, "vkh_free: pointer was not allocated by this context");
  size_t sz = it->second;
  live_.erase(it);
  live_bytes_ -= sz;
  free_[sz].push_back(p);
  cached_bytes_ += sz;
}

void Allocator::empty_cache() {
  std::lock_guard<std::mutex> lk(mu_);
  for (auto& kv : free_)
    for (void* q : kv.second) (void)hipFree(q);
  free_.clear();
  cached_bytes_ = 0;
}

Allocator::~Allocator() {
  empty_cache();
  for (auto& kv : live_) (void)hipFree(kv.first);
}

Workspace::Workspace(vkh_gpu* g, size_t nbytes) : g_(g) {
  if (nbytes == 0) nbytes = 1;
  if (g->capturing) {
    // A captured graph uses this memory on every replay: take it from the
    // graph's arena (nodes of one graph run in stream order, so a plain bump
    // allocation never overlaps live data).
    const size_t need = ceil_div(nbytes, 256) * 256;
    if (!g->arena || g->arena_used + need > g->arena_size) {
      g->arena_hint = g->arena_used + need;
      fail(VKH_OUT_OF_MEMORY, "graph workspace arena exhausted (" + std::to_string(g->arena_size) +
                                  " bytes); capture again to reserve a larger arena");
    }
    p_ = g->arena + g->arena_used;
    g->arena_used += need;
    graph_owned_ = true;
  } else if (g->alloc) {
    p_ = g->alloc->allocate(nbytes);
  } else {
    VKH_HIP(hipMallocAsync(&p_, nbytes, g->stream));
  }
}

Workspace::~Workspace() {
  if (!p_ || graph_owned_) return;
  if (g_->alloc) {
    try { g_->alloc->release(p_); } catch (...) {}
  } else {
    (void)hipFreeAsync(p_, g_->stream);
  }
}

// ---- fill kernels ------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_fill(T* __restrict__ dst, T value, uint64_t n) {
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] = value;
}

template <typename T>
static void fill(vkh_gpu* g, T* dst, T value, uint64_t n) {
  if (n == 0) return;
  VKH_REQUIRE(dst != nullptr, "fill: null pointer");
  unsigned grid = stream_grid(g, n, 256);
  hipLaunchKernelGGL(k_fill<T>, dim3(grid), dim3(256), 0, g->stream, dst, value, n);
  check_launch("fill");
}

static void init_props(vkh_gpu* g) {
  hipDeviceProp_t prop;
  VKH_HIP(hipGetDeviceProperties(&prop, g->device));
  g->num_cus = prop.multiProcessorCount;
  g->lds_bytes = (int)prop.sharedMemPerBlock;
  g->wave_size = prop.warpSize;
}

// ---- pinned host blocks ----------------------------------------------------------
// Page-locked host memory (hipHostMalloc) for the host mirrors of arrays, so
// that uploads and downloads are single DMA transfers instead of copies staged
// through the runtime's own buffers.  Pinning is slow (the pages are locked one
// by one), so freed blocks are cached by size class, up to HOST_CACHE_MAX
// bytes; live pinned memory is capped at HOST_LIVE_MAX (the caller falls back
// to ordinary memory beyond it).
constexpr size_t HOST_CACHE_MAX = size_t(4) << 30;
constexpr size_t HOST_LIVE_MAX = size_t(32) << 30;

struct HostPool {
  std::mutex mu;
  std::unordered_map<size_t, std::vector<void*>> free_;
  std::unordered_map<void*, size_t> live_;
  size_t live_bytes = 0, cached_bytes = 0;

  static size_t round(size_t n) {
    const size_t page = 4096, large = size_t(2) << 20;
    if (n == 0) n = 1;
    return n < (size_t(1) << 20) ? ceil_div(n, page) * page : ceil_div(n, large) * large;
  }
};

static HostPool& host_pool() {
  static HostPool* pool = new HostPool;  // kept until the process ends
  return *pool;
}

}  // namespace vkh

using namespace vkh;

extern "C" {

const char* vkh_last_error(void) { return g_last_error.c_str(); }

int vkh_abi_version(int* out) {
  return guarded([&] { VKH_REQUIRE(out, "null out"); *out = 1; });
}

int vkh_device_count(int* out) {
  return guarded([&] {
    VKH_REQUIRE(out, "null out");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) { (void)hipGetLastError(); n = 0; }
    *out = n;
  });
}

int vkh_gpu_create(int idx, float priority, vkh_gpu** out) {
  return guarded([&] {
    VKH_REQUIRE(out, "null out");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) { (void)hipGetLastError(); fail(VKH_NO_DEVICE, "no HIP device is visible"); }
    if (idx < 0 || idx >= n) fail(VKH_NO_DEVICE, "GPU index " + std::to_string(idx) + " out of range (" + std::to_string(n) + " devices)");
    auto* g = new vkh_gpu();
    g->device = idx;
    g->priority = priority;
    try {
      VKH_HIP(hipSetDevice(idx));
      // Default priority on purpose: streams of differing priorities are not used.
      VKH_HIP(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
      g->own_stream = true;
      g->alloc = new Allocator(idx);
      init_props(g);
    } catch (...) {
      if (g->stream) (void)hipStreamDestroy(g->stream);
      delete g->alloc;
      delete g;
      throw;
    }
    *out = g;
  });
}

int vkh_gpu_create_on_stream(int idx, void* stream, vkh_gpu** out) {
  return guarded([&] {
    VKH_REQUIRE(out, "null out");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) { (void)hipGetLastError(); fail(VKH_NO_DEVICE, "no HIP device is visible"); }
    if (idx < 0 || idx >= n) fail(VKH_NO_DEVICE, "GPU index out of range");
    auto* g = new vkh_gpu();
    g->device = idx;
    g->stream = (hipStream_t)stream;
    g->own_stream = false;
    try {
      VKH_HIP(hipSetDevice(idx));
      init_props(g);
      // workspaces and packing areas of this context are reused in the order
      // of the caller's stream, which is the only stream the context uses
      g->alloc = new Allocator(idx);
    } catch (...) { delete g->alloc; delete g; throw; }
    *out = g;
  });
}

int vkh_gpu_destroy(vkh_gpu* g) {
  return guarded([&] {
    if (!g) return;
    VKH_HIP(hipSetDevice(g->device));
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    delete g->alloc;
    if (g->own_stream && g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
  });
}

int vkh_gpu_wait(vkh_gpu* g) {
  return guarded([&] { use(g); VKH_HIP(hipStreamSynchronize(g->stream)); });
}

int vkh_gpu_can_subgroup_arithmetic(vkh_gpu* g, int* out) {
  return guarded([&] { VKH_REQUIRE(g && out, "null argument"); *out = 1; });
}

int vkh_gpu_get_stream(vkh_gpu* g, void** out) {
  return guarded([&] { VKH_REQUIRE(g && out, "null argument"); *out = (void*)g->stream; });
}

int vkh_gpu_info(vkh_gpu* g, int* num_cus, int* lds_bytes, int* wave_size) {
  return guarded([&] {
    VKH_REQUIRE(g, "null GPU handle");
    if (num_cus) *num_cus = g->num_cus;
    if (lds_bytes) *lds_bytes = g->lds_bytes;
    if (wave_size) *wave_size = g->wave_size;
  });
}

int vkh_alloc(vkh_gpu* g, size_t nbytes, void** out) {
  return guarded([&] {
    use(g);
    VKH_REQUIRE(out, "null out");
    VKH_REQUIRE(g->alloc, "context on an external stream has no allocator");
    VKH_REQUIRE(!g->capturing, "device allocation while a graph is being captured");
    *out = g->alloc->allocate(nbytes);
  });
}

int vkh_free(vkh_gpu* g, void* p) {
  return guarded([&] {
    if (!p) return;
    use(g);
    VKH_REQUIRE(g->alloc, "context on an external stream has no allocator");
    g->alloc->release(p);
  });
}

int vkh_empty_cache(vkh_gpu* g) {
  return guarded([&] {
    use(g);
    if (!g->alloc) return;
    VKH_HIP(hipStreamSynchronize(g->stream));
    g->alloc->empty_cache();
  });
}

int vkh_memory_stats(vkh_gpu* g, size_t* live, size_t* cached) {
  return guarded([&] {
    VKH_REQUIRE(g, "null GPU handle");
    if (live) *live = g->alloc ? g->alloc->live_bytes() : 0;
    if (cached) *cached = g->alloc ? g->alloc->cached_bytes() : 0;
  });
}

int vkh_upload(vkh_gpu* g, void* dst, const void* src, size_t nbytes) {
  return guarded([&] {
    use(g);
    // A captured copy would read the host buffer again at every replay, long
    // after the caller may have reused it.
    VKH_REQUIRE(!g->capturing, "upload: not allowed during graph capture");
    if (nbytes == 0) return;
    VKH_REQUIRE(dst && src, "upload: null pointer");
    VKH_HIP(hipMemcpyAsync(dst, src, nbytes, hipMemcpyHostToDevice, g->stream));
  });
}

int vkh_download(vkh_gpu* g, void* dst, const void* src, size_t nbytes) {
  return guarded([&] {
    use(g);
    // the copy is waited for on the host, which a capturing stream cannot do
    VKH_REQUIRE(!g->capturing, "download: not allowed during graph capture");
    if (nbytes == 0) return;
    VKH_REQUIRE(dst && src, "download: null pointer");
    VKH_HIP(hipMemcpyAsync(dst, src, nbytes, hipMemcpyDeviceToHost, g->stream));
    VKH_HIP(hipStreamSynchronize(g->stream));
  });
}

int vkh_host_alloc(size_t nbytes, void** out) {
  return guarded([&] {
    VKH_REQUIRE(out, "host_alloc: null out");
    *out = nullptr;
    HostPool& hp = host_pool();
    const size_t sz = HostPool::round(nbytes);
    std::lock_guard<std::mutex> lk(hp.mu);
    auto it = hp.free_.find(sz);
    if (it != hp.free_.end() && !it->second.empty()) {
      *out = it->second.back();
      it->second.pop_back();
      hp.cached_bytes -= sz;
    } else {
      if (hp.live_bytes + sz > HOST_LIVE_MAX) fail(VKH_OUT_OF_MEMORY, "host_alloc: pinned memory cap reached");
      void* p = nullptr;
      if (hipHostMalloc(&p, sz, hipHostMallocDefault) != hipSuccess || !p) {
        (void)hipGetLastError();
        fail(VKH_OUT_OF_MEMORY, "host_alloc: hipHostMalloc of " + std::to_string(sz) + " bytes failed");
      }
      *out = p;
    }
    hp.live_[*out] = sz;
    hp.live_bytes += sz;
  });
}

int vkh_host_free(void* p) {
  return guarded([&] {
    if (!p) return;
    HostPool& hp = host_pool();
    std::lock_guard<std::mutex> lk(hp.mu);
    auto it = hp.live_.find(p);
    VKH_REQUIRE(it != hp.live_.end(), "host_free: not a pinned block of this library");
    const size_t sz = it->second;
    hp.live_.erase(it);
    hp.live_bytes -= sz;
    if (hp.cached_bytes + sz <= HOST_CACHE_MAX) {
      hp.free_[sz].push_back(p);
      hp.cached_bytes += sz;
    } else {
      VKH_HIP(hipHostFree(p));
    }
  });
}

int vkh_host_empty_cache(void) {
  return guarded([&] {
    HostPool& hp = host_pool();
    std::lock_guard<std::mutex> lk(hp.mu);
    for (auto& kv : hp.free_)
      for (void* p : kv.second) VKH_HIP(hipHostFree(p));
    hp.free_.clear();
    hp.cached_bytes = 0;
  });
}

int vkh_host_stats(size_t* live, size_t* cached) {
  return guarded([&] {
    HostPool& hp = host_pool();
    std::lock_guard<std::mutex> lk(hp.mu);
    if (live) *live = hp.live_bytes;
    if (cached) *cached = hp.cached_bytes;
  });
}

int vkh_copy(vkh_gpu* g, void* dst, const void* src, size_t nbytes) {
  return guarded([&] {
    use(g);
    if (nbytes == 0) return;
    VKH_REQUIRE(dst && src, "copy: null pointer");
    VKH_HIP(hipMemcpyAsync(dst, src, nbytes, hipMemcpyDeviceToDevice, g->stream));
  });
}

int vkh_fill_f32(vkh_gpu* g, float* dst, float value, size_t n) {
  return guarded([&] { use(g); fill<float>(g, dst, value, n); });
}

int vkh_fill_u32(vkh_gpu* g, uint32_t* dst, uint32_t value, size_t n) {
  return guarded([&] { use(g); fill<uint32_t>(g, dst, value, n); });
}

int vkh_event_record(vkh_gpu* g, vkh_event** out) {
  return guarded([&] {
    use(g);
    VKH_REQUIRE(out, "null out");
    // a record on a capturing stream would become a node of the graph that
    // refers to this event after its owner has destroyed it
    VKH_REQUIRE(!g->capturing, "event_record: not allowed during graph capture");
    auto* e = new vkh_event();
    e->device = g->device;
    hipError_t r = hipEventCreateWithFlags(&e->ev, hipEventDisableTiming);
    if (r != hipSuccess) { delete e; check_hip(r, "hipEventCreateWithFlags"); }
    r = hipEventRecord(e->ev, g->stream);
    if (r != hipSuccess) { (void)hipEventDestroy(e->ev); delete e; check_hip(r, "hipEventRecord"); }
    *out = e;
  });
}

int vkh_timer_record(vkh_gpu* g, vkh_event** out) {
  return guarded([&] {
    use(g);
    VKH_REQUIRE(out, "null out");
    VKH_REQUIRE(!g->capturing, "timer_record: not allowed during graph capture");
    auto* e = new vkh_event();
    e->device = g->device;
    hipError_t r = hipEventCreate(&e->ev);
    if (r != hipSuccess) { delete e; check_hip(r, "hipEventCreate"); }
    r = hipEventRecord(e->ev, g->stream);
    if (r != hipSuccess) { (void)hipEventDestroy(e->ev); delete e; check_hip(r, "hipEventRecord"); }
    *out = e;
  });
}

int vkh_timer_next_kernel(vkh_gpu* g, vkh_event** start, vkh_event** stop) {
  return guarded([&] {
    use(g);
    VKH_REQUIRE(start && stop, "null out");
    VKH_REQUIRE(!g->capturing, "timer_next_kernel: not allowed during graph capture");
    vkh_event* ev[2] = {nullptr, nullptr};
    for (auto& e : ev) {
      e = new vkh_event();
      e->device = g->device;
      const hipError_t r = hipEventCreate(&e->ev);
      if (r != hipSuccess) {
        if (ev[0] && ev[0] != e) { (void)hipEventDestroy(ev[0]->ev); delete ev[0]; }
        delete e;
        check_hip(r, "hipEventCreate");
      }
    }
    g->ktime_start = ev[0]->ev;
    g->ktime_stop = ev[1]->ev;
    *start = ev[0];
    *stop = ev[1];
  });
}

int vkh_event_elapsed_ms(vkh_event* a, vkh_event* b, float* ms) {
  return guarded([&] {
    VKH_REQUIRE(a && b && ms, "null argument");
    VKH_HIP(hipSetDevice(a->device));
    VKH_HIP(hipEventSynchronize(b->ev));
    VKH_HIP(hipEventElapsedTime(ms, a->ev, b->ev));
  });
}

int vkh_event_wait(vkh_event* e, uint64_t timeout_ns) {
  return guarded([&] {
    VKH_REQUIRE(e, "null event");
    VKH_HIP(hipSetDevice(e->device));
    if (timeout_ns == UINT64_MAX) {
      hipError_t r = hipEventSynchronize(e->ev);
      if (r != hipSuccess) fail(VKH_ERROR, std::string("Error at Command Wait: ") + hipGetErrorString(r));
      return;
    }
    auto deadline = std::chrono::steady_clock::now() + std::chrono::nanoseconds(timeout_ns);
    for (;;) {
      hipError_t r = hipEventQuery(e->ev);
      if (r == hipSuccess) return;
      if (r != hipErrorNotReady) fail(VKH_ERROR, std::string("Error at Command Wait: ") + hipGetErrorString(r));
      if (std::chrono::steady_clock::now() >= deadline) fail(VKH_TIMEOUT, "Timeout at Command Wait");
      std::this_thread::yield();
    }
  });
}

int vkh_event_query(vkh_event* e, int* done) {
  return guarded([&] {
    VKH_REQUIRE(e && done, "null argument");
    VKH_HIP(hipSetDevice(e->device));
    hipError_t r = hipEventQuery(e->ev);
    if (r == hipSuccess) *done = 1;
    else if (r == hipErrorNotReady) *done = 0;
    else check_hip(r, "hipEventQuery");
  });
}

int vkh_event_destroy(vkh_event* e) {
  return guarded([&] {
    if (!e) return;
    (void)hipSetDevice(e->device);
    (void)hipEventDestroy(e->ev);
    delete e;
  });
}

int vkh_graph_begin(vkh_gpu* g) {
  return guarded([&] {
    use(g);
    VKH_REQUIRE(!g->capturing, "graph capture already active on this context");
    VKH_REQUIRE(g->alloc, "graph capture needs a context with its own allocator");
    // reserve the arena (workspaces, then the GEMM packing region) before the
    // stream starts capturing
    const size_t ws = ceil_div(std::max<size_t>(size_t(64) << 20, g->arena_hint), 256) * 256;
    const size_t pk = ceil_div(g->graph_pack_hint, 256) * 256;
    g->arena = (char*)g->alloc->allocate(ws + pk);
    g->arena_size = ws;
    g->arena_used = 0;
    g->graph_pack = pk ? g->arena + ws : nullptr;
    g->graph_pack_bytes = pk;
    hipError_t r = hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal);
    if (r != hipSuccess) {
      g->alloc->release(g->arena);
      g->arena = nullptr;
      g->graph_pack = nullptr;
      g->graph_pack_bytes = 0;
      check_hip(r, "hipStreamBeginCapture");
    }
    g->capturing = true;
  });
}

int vkh_graph_end(vkh_gpu* g, void** exec) {
  return guarded([&] {
    VKH_REQUIRE(g, "null GPU handle");
    VKH_REQUIRE(g->capturing, "no graph capture active on this context");
    // End the capture before anything that may fail, so the stream is usable
    // again whatever happened while capturing.
    g->capturing = false;
    hipGraph_t graph = nullptr;
    hipError_t r = hipStreamEndCapture(g->stream, &graph);
    auto* gr = new vkh::Graph();
    gr->gpu = g;
    gr->arena = g->arena;
    g->arena = nullptr;
    g->arena_size = g->arena_used = 0;
    g->graph_pack = nullptr;
    g->graph_pack_bytes = 0;
    if (r == hipSuccess) {
      r = hipGraphInstantiate(&gr->exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
    } else if (graph) {
      (void)hipGraphDestroy(graph);
    }
    if (r != hipSuccess || !exec) {
      if (gr->exec) (void)hipGraphExecDestroy(gr->exec);
      g->alloc->release(gr->arena);
      delete gr;
      VKH_REQUIRE(exec, "null out");
      check_hip(r, "graph capture");
    }
    *exec = (void*)gr;
  });
}

int vkh_graph_launch(vkh_gpu* g, void* exec) {
  return guarded([&] {
    use(g);
    VKH_REQUIRE(exec, "null graph");
    auto* gr = (vkh::Graph*)exec;
    VKH_REQUIRE(gr->gpu == g, "graph belongs to another context");
    VKH_HIP(hipGraphLaunch(gr->exec, g->stream));
  });
}

int vkh_graph_destroy(void* exec) {
  return guarded([&] {
    if (!exec) return;
    auto* gr = (vkh::Graph*)exec;
    use(gr->gpu);
    // the release waits for the stream, which a capture cannot do
    VKH_REQUIRE(!gr->gpu->capturing, "graph_destroy: not allowed during graph capture");
    VKH_HIP(hipStreamSynchronize(gr->gpu->stream));
    if (gr->exec) VKH_HIP(hipGraphExecDestroy(gr->exec));
    if (gr->arena) gr->gpu->alloc->release(gr->arena);
    delete gr;
  });
}

}  // extern "C"
