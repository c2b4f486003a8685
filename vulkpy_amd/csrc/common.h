// Internal helpers shared by the HIP translation units of libvulkpy_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstdint>
#include <map>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/vulkpy_hip.h"
#include "bmath.h"

namespace vkh {

// ---- errors ---------------------------------------------------------------
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_error(const std::string& msg);
void clear_error();

[[noreturn]] inline void fail(int code, const std::string& msg) {
  throw Error(code, msg);
}

inline void check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    int code = (e == hipErrorOutOfMemory) ? VKH_OUT_OF_MEMORY : VKH_ERROR;
    fail(code, std::string(what) + ": " + hipGetErrorString(e));
  }
}

#define VKH_HIP(call) ::vkh::check_hip((call), #call)

// Every exported function body runs inside guarded(): exceptions become codes.
template <typename F>
int guarded(F&& f) {
  try {
    clear_error();
    f();
    return VKH_OK;
  } catch (const Error& e) {
    set_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_error("host out of memory");
    return VKH_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    set_error(e.what());
    return VKH_ERROR;
  }
}

#define VKH_REQUIRE(cond, msg) \
  do { if (!(cond)) ::vkh::fail(VKH_INVALID, (msg)); } while (0)

// ---- stream-ordered caching allocator -------------------------------------
// Blocks are rounded to a size class and kept in per-class free lists after
// vkh_free.  Reuse is safe without synchronisation because every user of a
// context enqueues on the context's single in-order stream.
class Allocator {
 public:
  explicit Allocator(int device) : device_(device) {}
  int device() const { return device_; }
  ~Allocator();
  void* allocate(size_t nbytes);
  void release(void* p);
  void empty_cache();
  size_t live_bytes() const { return live_bytes_; }
  size_t cached_bytes() const { return cached_bytes_; }
  static size_t round_size(size_t nbytes);

 private:
  int device_;
  std::mutex mu_;
  std::unordered_map<void*, size_t> live_;
  std::unordered_map<size_t, std::vector<void*>> free_;
  size_t live_bytes_ = 0;
  size_t cached_bytes_ = 0;
};

}  // namespace vkh

// ---- context ---------------------------------------------------------------
struct vkh_gpu {
  int device = 0;
  float priority = 0.0f;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  vkh::Allocator* alloc = nullptr;  // null for contexts on external streams
  int num_cus = 0;
  int lds_bytes = 0;
  int wave_size = 64;
  bool capturing = false;        // stream capture in progress (vkh_graph_begin)
  // Workspaces of kernels captured into a graph are bump-allocated from this
  // arena, reserved before the capture begins (no allocation is legal while
  // the stream captures) and owned by the resulting graph.
  char* arena = nullptr;
  size_t arena_size = 0, arena_used = 0, arena_hint = 0;
  // grow-only BF16 packing area of the matrix-core GEMMs (gemm.hip) for
  // eager calls
  void* pack = nullptr;
  size_t pack_bytes = 0;
  // packing area of GEMMs captured into a graph: a region of that graph's
  // arena shared by all of its GEMM nodes (they run in stream order), sized
  // from graph_pack_hint (vkh_gemm_reserve) when the capture begins
  void* graph_pack = nullptr;
  size_t graph_pack_bytes = 0, graph_pack_hint = 0;
  // GEMM kernel configuration per (mode, trans_a, trans_b, M, N, K), chosen
  // by vkh_gemm_autotune or vkh_gemm_set_config (gemm.hip)
  std::map<std::array<uint64_t, 6>, int> gemm_cfg;
  // reduction kernel configuration per (family, prev, axis, post), chosen by
  // vkh_reduce_autotune or vkh_reduce_set_config (reduce.hip)
  std::map<std::array<uint64_t, 4>, int> red_cfg;
  // events armed by vkh_timer_next_kernel: the next elementwise kernel's
  // dispatch records them at its own start and end (elementwise.hip)
  hipEvent_t ktime_start = nullptr, ktime_stop = nullptr;
};

struct vkh_event {
  hipEvent_t ev = nullptr;
  int device = 0;
};

namespace vkh {

// Makes the context's device current for the calling thread.
// Column-sum order shared with vkh_bias_grad (reduce.hip): rows per chunk of
// a [rows, cols] operand, and the in-order combine of chunk partials.
uint64_t bias_grad_chunk(vkh_gpu* g, uint64_t rows, uint64_t cols, bool vec);
void sum_chunks(vkh_gpu* g, const float* part, float* out, uint64_t m, uint64_t nchunk);
void sum_chunks_from_zero(vkh_gpu* g, const float* part, float* out, uint64_t m, uint64_t nchunk);

inline void use(vkh_gpu* g) {
  VKH_REQUIRE(g != nullptr, "null GPU handle");
  VKH_HIP(hipSetDevice(g->device));
}

// Temporary device buffer taken from the context allocator (or hipMalloc for
// contexts without one) and returned when the holder goes out of scope; the
// return is stream ordered, so kernels enqueued before release may still use
// it.
class Workspace {
 public:
  Workspace(vkh_gpu* g, size_t nbytes);
  ~Workspace();
  void* get() const { return p_; }
  Workspace(const Workspace&) = delete;
  Workspace& operator=(const Workspace&) = delete;

 private:
  vkh_gpu* g_;
  void* p_ = nullptr;
  bool graph_owned_ = false;
};

// Executable graph plus the workspace arena its kernels reference.
struct Graph {
  vkh_gpu* gpu = nullptr;
  hipGraphExec_t exec = nullptr;
  void* arena = nullptr;
};

inline void check_launch(const char* name) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) fail(VKH_ERROR, std::string("launch of ") + name + ": " + hipGetErrorString(e));
}

inline uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

// Grid size for a grid-stride streaming kernel: enough blocks to cover the
// work, capped at `per_cu` resident blocks per compute unit.
inline unsigned stream_grid(const vkh_gpu* g, uint64_t work_items, unsigned block,
                            unsigned per_cu = 8) {
  uint64_t need = ceil_div(work_items, block);
  uint64_t cap = (uint64_t)(g->num_cus > 0 ? g->num_cus : 256) * per_cu;
  if (need < 1) need = 1;
  return (unsigned)(need < cap ? need : cap);
}

// ---- streaming memory access ----------------------------------------------
// Nontemporal 128-bit loads / stores for data touched once (operands larger
// than the caches): measured 6.5-6.9 TB/s on MI355X against 5.8-6.2 TB/s for
// ordinary accesses in the same one-float4-per-thread kernels.
typedef float vkh_f4v __attribute__((ext_vector_type(4)));

// FP32 -> BF16 bits: round to nearest even, NaN kept NaN (quiet), and a
// nonzero value that would round to zero (|x| < 2^-134) keeps the smallest
// BF16 magnitude of its sign instead.  So a BF16 copy has the sign, the
// zero-ness and the NaN-ness of its value, and a ReLU' factor read from it
// equals the one read from the FP32 value.
__device__ __forceinline__ uint16_t f2bf(float x) {
  const uint32_t u = __float_as_uint(x);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  uint32_t b = (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
  if ((b & 0x7fffu) == 0 && (u & 0x7fffffffu) != 0) b |= 1u;
  return (uint16_t)b;
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ float4 ld_stream(const float4* p) {
  const vkh_f4v v = __builtin_nontemporal_load(reinterpret_cast<const vkh_f4v*>(p));
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void st_stream(float4* p, const float4& v) {
  vkh_f4v w = {v.x, v.y, v.z, v.w};
  __builtin_nontemporal_store(w, reinterpret_cast<vkh_f4v*>(p));
}
template <typename T>
__device__ __forceinline__ void st_stream1(T* p, T v) {
  __builtin_nontemporal_store(v, p);
}

// Fills p[0..n) with hashed values in [-0.5, 0.5): the scratch operands of
// the autotuning runs (gemm.hip, block of 256 threads).
__global__ void __launch_bounds__(256) k_fill_noise(float* p, uint64_t n, uint32_t seed);

// Grid for a one-work-item-per-thread kernel (with a grid-stride loop as a
// guard for very large arrays).
inline unsigned full_grid(uint64_t work_items, unsigned block) {
  uint64_t need = ceil_div(work_items, block);
  if (need < 1) need = 1;
  const uint64_t cap = uint64_t(1) << 24;
  return (unsigned)(need < cap ? need : cap);
}

// ---- device math shared by elementwise, reduction and nn kernels ----------
// max / min follow numpy.maximum / numpy.minimum (NaN propagates, ties keep a).
__device__ __forceinline__ float np_max(float a, float b) {
  return (a >= b || a != a) ? a : b;
}
__device__ __forceinline__ float np_min(float a, float b) {
  return (a <= b || a != a) ? a : b;
}

// Transcendentals are evaluated in FP64 and rounded once to FP32, which keeps
// them within one rounding of the exact value (the FP64 oracle).
__device__ __forceinline__ float dexp(float x) { return (float)exp((double)x); }
__device__ __forceinline__ float dlog(float x) { return (float)log((double)x); }

__device__ __forceinline__ float binop_apply(int op, float a, float b) {
  switch (op) {
    case VKH_ADD: return a + b;
    case VKH_SUB: return a - b;
    case VKH_MUL: return a * b;
    case VKH_DIV: return a / b;
    case VKH_MAX: return np_max(a, b);
    case VKH_MIN: return np_min(a, b);
    case VKH_POW: return pow_f32(a, b);
    case VKH_RSUB: return b - a;
    case VKH_RDIV: return b / a;
    case VKH_RPOW: return pow_f32(b, a);
  }
  return 0.0f;
}

template <int OP>
__device__ __forceinline__ float binop(float a, float b) {
  if constexpr (OP == VKH_ADD) return a + b;
  else if constexpr (OP == VKH_SUB) return a - b;
  else if constexpr (OP == VKH_MUL) return a * b;
  else if constexpr (OP == VKH_DIV) return a / b;
  else if constexpr (OP == VKH_MAX) return np_max(a, b);
  else if constexpr (OP == VKH_MIN) return np_min(a, b);
  else if constexpr (OP == VKH_POW) return pow_f32(a, b);
  else if constexpr (OP == VKH_RSUB) return b - a;
  else if constexpr (OP == VKH_RDIV) return b / a;
  else return pow_f32(b, a);
}

// Dispatch a runtime binary op code to a template parameter.
template <template <int> class K, typename... Args>
void dispatch_binop(int op, Args&&... args) {
  switch (op) {
    case VKH_ADD: K<VKH_ADD>::run(args...); break;
    case VKH_SUB: K<VKH_SUB>::run(args...); break;
    case VKH_MUL: K<VKH_MUL>::run(args...); break;
    case VKH_DIV: K<VKH_DIV>::run(args...); break;
    case VKH_MAX: K<VKH_MAX>::run(args...); break;
    case VKH_MIN: K<VKH_MIN>::run(args...); break;
    case VKH_POW: K<VKH_POW>::run(args...); break;
    case VKH_RSUB: K<VKH_RSUB>::run(args...); break;
    case VKH_RDIV: K<VKH_RDIV>::run(args...); break;
    case VKH_RPOW: K<VKH_RPOW>::run(args...); break;
    default: fail(VKH_INVALID, "unknown binary op " + std::to_string(op));
  }
}

}  // namespace vkh
