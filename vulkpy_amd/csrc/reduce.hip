// Reductions: full (F11/F12), axis (F13) and axis + rebroadcast (F14).
//
// The reference reduces an axis with one thread per output element looping
// sequentially over k (vulkpy/shader/sum_axis.comp:20-32) and the whole array
// with a tree of ceil(n/64) passes (vulkpy/vkarray.py:1246-1274).  Here:
//  * rows (post == 1): one block (long rows), one wave (medium) or one thread
//    (short rows) per output, 128-bit loads, wave shuffle + LDS tree;
//  * strided axes (post > 1): one thread per output column group (float4 when
//    post % 4 == 0) looping sequentially over k -- the reference order -- and,
//    when that leaves the GPU under-occupied, k is split into chunks whose
//    partials are combined in chunk order by a second kernel;
//  * full reductions: grid-stride FP64 accumulation per thread, FP64 block
//    tree, FP64 cross-block combine, one rounding to FP32 at the end.
#include "common.h"

namespace vkh {

template <int OP, typename T>
struct Red;
template <typename T>
struct Red<VKH_RSUM, T> {
  static __device__ __forceinline__ T init() { return T(0); }
  static __device__ __forceinline__ T f(T a, T b) { return a + b; }
};
template <typename T>
struct Red<VKH_RPROD, T> {
  static __device__ __forceinline__ T init() { return T(1); }
  static __device__ __forceinline__ T f(T a, T b) { return a * b; }
};
template <typename T>
struct Red<VKH_RMAX, T> {
  static __device__ __forceinline__ T init() { return -__builtin_huge_valf(); }
  static __device__ __forceinline__ T f(T a, T b) { return (a >= b || a != a) ? a : b; }
};
template <typename T>
struct Red<VKH_RMIN, T> {
  static __device__ __forceinline__ T init() { return __builtin_huge_valf(); }
  static __device__ __forceinline__ T f(T a, T b) { return (a <= b || a != a) ? a : b; }
};

template <int OP, typename T>
__device__ __forceinline__ T wave_reduce(T v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = Red<OP, T>::f(v, __shfl_xor(v, off, 64));
  return v;
}

// Block-wide reduction; every thread gets the result.  BLOCK is a multiple of 64.
template <int OP, typename T, int BLOCK>
__device__ __forceinline__ T block_reduce(T v, T* lds) {
  v = wave_reduce<OP, T>(v);
  constexpr int NW = BLOCK / 64;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (NW == 1) return v;
  if (lane == 0) lds[w] = v;
  __syncthreads();
  T r = lds[0];
#pragma unroll
  for (int i = 1; i < NW; ++i) r = Red<OP, T>::f(r, lds[i]);
  __syncthreads();
  return r;
}

__device__ __forceinline__ float epilogue(float v, float mul, float dv) {
  if (mul != 1.0f) v = v * mul;
  if (dv != 1.0f) v = v / dv;
  return v;
}

// ---- rows: post == 1, reduce each contiguous row of length L -----------------
constexpr int ROW_BLOCK = 256;

// one block per row (long rows)
// U: 128-bit loads in flight per thread (4 or 8); each thread still combines
// its elements in index order, so U never changes the result
template <int OP, bool VEC, int U>
__global__ void __launch_bounds__(ROW_BLOCK)
k_rows_block(const float* a, float* out, uint64_t rows, uint64_t L, int rebroadcast, float mul, float dv) {
  __shared__ float lds[ROW_BLOCK / 64];
  for (uint64_t r = blockIdx.x; r < rows; r += gridDim.x) {
    const float* row = a + r * L;
    float acc = Red<OP, float>::init();
    if (VEC) {
      using R = Red<OP, float>;
      const float4* row4 = reinterpret_cast<const float4*>(row);
      const uint64_t L4 = L / 4;
      uint64_t i = threadIdx.x;
      // U loads in flight, combined in the same order as the plain loop
      for (; i + (U - 1) * ROW_BLOCK < L4; i += U * ROW_BLOCK) {
        float4 v[U];
#pragma unroll
        for (int q = 0; q < U; ++q) v[q] = ld_stream(row4 + i + q * ROW_BLOCK);
#pragma unroll
        for (int q = 0; q < U; ++q) acc = R::f(acc, R::f(R::f(v[q].x, v[q].y), R::f(v[q].z, v[q].w)));
      }
      for (; i < L4; i += ROW_BLOCK) {
        const float4 v = ld_stream(row4 + i);
        acc = R::f(acc, R::f(R::f(v.x, v.y), R::f(v.z, v.w)));
      }
    } else {
      for (uint64_t i = threadIdx.x; i < L; i += ROW_BLOCK) acc = Red<OP, float>::f(acc, row[i]);
    }
    float res = epilogue(block_reduce<OP, float, ROW_BLOCK>(acc, lds), mul, dv);
    if (!rebroadcast) {
      if (threadIdx.x == 0) out[r] = res;
    } else {
      float* orow = out + r * L;
      if (VEC) {
        float4* o4 = reinterpret_cast<float4*>(orow);
        for (uint64_t i = threadIdx.x; i < L / 4; i += ROW_BLOCK) o4[i] = make_float4(res, res, res, res);
      } else {
        for (uint64_t i = threadIdx.x; i < L; i += ROW_BLOCK) orow[i] = res;
      }
    }
  }
}

// one wave per row (medium rows); 4 rows per block
template <int OP>
__global__ void __launch_bounds__(256)
k_rows_wave(const float* a, float* out, uint64_t rows, uint64_t L, int rebroadcast, float mul, float dv) {
  const int lane = threadIdx.x & 63;
  const uint64_t nw_total = (uint64_t)gridDim.x * 4;
  for (uint64_t r = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6); r < rows; r += nw_total) {
    const float* row = a + r * L;
    float acc = Red<OP, float>::init();
    for (uint64_t i = lane; i < L; i += 64) acc = Red<OP, float>::f(acc, row[i]);
    float res = epilogue(wave_reduce<OP, float>(acc), mul, dv);
    if (!rebroadcast) {
      if (lane == 0) out[r] = res;
    } else {
      for (uint64_t i = lane; i < L; i += 64) out[r * L + i] = res;
    }
  }
}

// one thread per row (short rows), sequential over the row like the reference
template <int OP>
__global__ void __launch_bounds__(256)
k_rows_thread(const float* a, float* out, uint64_t rows, uint64_t L, int rebroadcast, float mul, float dv) {
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t r = (uint64_t)blockIdx.x * 256 + threadIdx.x; r < rows; r += stride) {
    const float* row = a + r * L;
    float acc = row[0];
    for (uint64_t i = 1; i < L; ++i) acc = Red<OP, float>::f(acc, row[i]);
    float res = epilogue(acc, mul, dv);
    if (!rebroadcast) out[r] = res;
    else for (uint64_t i = 0; i < L; ++i) out[r * L + i] = res;
  }
}

// ---- strided axes: post > 1 ------------------------------------------------------
// a viewed as [prev, K, post]; thread owns VW consecutive columns j of one p
// and reduces k in [k0, k1) sequentially.  Result goes to dst[(c * prev + p) *
// post + j] where c is the chunk index (blockIdx.y).  U rows are loaded at a
// time (VW == 4) and still combined in row order, so U never changes the result.
template <int OP, int VW, int U>
__global__ void __launch_bounds__(256)
k_cols(const float* a, float* dst, uint64_t prev, uint64_t K, uint64_t post, uint64_t kchunk,
       int final_pass, float mul, float dv) {
  const uint64_t groups = post / VW;
  const uint64_t total = prev * groups;
  const uint64_t c = blockIdx.y;
  const uint64_t k0 = c * kchunk;
  const uint64_t k1 = (k0 + kchunk < K) ? k0 + kchunk : K;
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x; t < total; t += stride) {
    const uint64_t p = t / groups, j = (t - p * groups) * VW;
    const float* base = a + (p * K) * post + j;
    if (VW == 4) {
      using R = Red<OP, float>;
      auto comb = [](float4& acc, const float4& v) {
        acc.x = R::f(acc.x, v.x);
        acc.y = R::f(acc.y, v.y);
        acc.z = R::f(acc.z, v.z);
        acc.w = R::f(acc.w, v.w);
      };
      float4 acc = ld_stream(reinterpret_cast<const float4*>(base + k0 * post));
      uint64_t k = k0 + 1;
      // U rows in flight, still combined in row order
      for (; k + (U - 1) < k1; k += U) {
        float4 v[U];
#pragma unroll
        for (int q = 0; q < U; ++q) v[q] = ld_stream(reinterpret_cast<const float4*>(base + (k + q) * post));
#pragma unroll
        for (int q = 0; q < U; ++q) comb(acc, v[q]);
      }
      for (; k < k1; ++k) comb(acc, ld_stream(reinterpret_cast<const float4*>(base + k * post)));
      if (final_pass) {
        acc.x = epilogue(acc.x, mul, dv); acc.y = epilogue(acc.y, mul, dv);
        acc.z = epilogue(acc.z, mul, dv); acc.w = epilogue(acc.w, mul, dv);
      }
      *reinterpret_cast<float4*>(dst + (c * prev + p) * post + j) = acc;
    } else {
      float acc = base[k0 * post];
      for (uint64_t k = k0 + 1; k < k1; ++k) acc = Red<OP, float>::f(acc, base[k * post]);
      if (final_pass) acc = epilogue(acc, mul, dv);
      dst[(c * prev + p) * post + j] = acc;
    }
  }
}

// combine nchunk partials [nchunk, m] in chunk order into out[m]
template <int OP>
__global__ void __launch_bounds__(256)
k_combine_chunks(const float* part, float* out, uint64_t m, uint64_t nchunk, float mul, float dv) {
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < m; i += stride) {
    float acc = part[i];
    uint64_t c = 1;
    for (; c + 8 <= nchunk; c += 8) {  // eight loads in flight, combined in chunk order
      float v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = part[(c + q) * m + i];
#pragma unroll
      for (int q = 0; q < 8; ++q) acc = Red<OP, float>::f(acc, v[q]);
    }
    for (; c < nchunk; ++c) acc = Red<OP, float>::f(acc, part[c * m + i]);
    out[i] = epilogue(acc, mul, dv);
  }
}

// red [prev, post] -> out [prev, K, post] (rebroadcast of a strided axis)
__global__ void __launch_bounds__(256)
k_rebroadcast_cols(const float* red, float* out, uint64_t prev, uint64_t K, uint64_t post) {
  const uint64_t total = prev * K * post;
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x; t < total; t += stride) {
    const uint64_t j = t % post, p = t / (K * post);
    out[t] = red[p * post + j];
  }
}

// ---- full reduction ----------------------------------------------------------------
constexpr int ALL_BLOCK = 256;

// U (4 or 8) 128-bit loads in flight per thread; a group of eight is combined
// as two consecutive groups of four, so for a given grid U never changes the
// result
template <int OP, int U>
__global__ void __launch_bounds__(ALL_BLOCK)
k_all_partial(const float* a, uint64_t n, double* part) {
  __shared__ double lds[ALL_BLOCK / 64];
  double acc = Red<OP, double>::init();
  const uint64_t stride = (uint64_t)gridDim.x * ALL_BLOCK;
  const uint64_t n4 = ((uintptr_t)a & 15u) == 0 ? n / 4 : 0;
  const float4* a4 = reinterpret_cast<const float4*>(a);
  using R = Red<OP, double>;
  auto quad = [](const float4& v) {
    return R::f(R::f((double)v.x, (double)v.y), R::f((double)v.z, (double)v.w));
  };
  uint64_t i = (uint64_t)blockIdx.x * ALL_BLOCK + threadIdx.x;
  if constexpr (U == 8) {
    for (; i + 7 * stride < n4; i += 8 * stride) {
      float4 v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = ld_stream(a4 + i + q * stride);
      acc = R::f(acc, R::f(R::f(quad(v[0]), quad(v[1])), R::f(quad(v[2]), quad(v[3]))));
      acc = R::f(acc, R::f(R::f(quad(v[4]), quad(v[5])), R::f(quad(v[6]), quad(v[7]))));
    }
  }
  for (; i + 3 * stride < n4; i += 4 * stride) {  // four loads in flight
    const float4 v0 = ld_stream(a4 + i), v1 = ld_stream(a4 + i + stride);
    const float4 v2 = ld_stream(a4 + i + 2 * stride), v3 = ld_stream(a4 + i + 3 * stride);
    acc = R::f(acc, R::f(R::f(quad(v0), quad(v1)), R::f(quad(v2), quad(v3))));
  }
  for (; i < n4; i += stride) acc = R::f(acc, quad(ld_stream(a4 + i)));
  for (uint64_t i = n4 * 4 + (uint64_t)blockIdx.x * ALL_BLOCK + threadIdx.x; i < n; i += stride)
    acc = Red<OP, double>::f(acc, (double)a[i]);
  double r = block_reduce<OP, double, ALL_BLOCK>(acc, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

template <int OP>
__global__ void __launch_bounds__(ALL_BLOCK)
k_all_final(const double* part, uint64_t nb, float* out, uint64_t n, int mean) {
  __shared__ double lds[ALL_BLOCK / 64];
  double acc = Red<OP, double>::init();
  for (uint64_t i = threadIdx.x; i < nb; i += ALL_BLOCK) acc = Red<OP, double>::f(acc, part[i]);
  double r = block_reduce<OP, double, ALL_BLOCK>(acc, lds);
  if (threadIdx.x == 0) out[0] = (float)(mean ? r / (double)n : r);
}

// ---- kernel configurations (vkh_reduce_autotune) ------------------------------
// unroll: 128-bit loads in flight per thread.  par: blocks per CU (full and row
// reductions: the grid cap) or target threads per CU (strided axes: decides
// how many chunks k is split into).  Index 0 of each family is the default.
struct RedCfg {
  int unroll, par;
};
static const RedCfg ALL_CFGS[] = {{4, 4}, {8, 4}, {4, 2}, {8, 2}, {4, 8}, {8, 8}, {4, 16}, {8, 16}};
static const RedCfg ROWS_CFGS[] = {{4, 16}, {8, 16}, {4, 8}, {8, 8}, {4, 32}, {8, 32}};
static const RedCfg COLS_CFGS[] = {{4, 512}, {8, 512}, {4, 256}, {8, 256},
                                   {4, 1024}, {8, 1024}, {4, 2048}, {8, 2048}};
constexpr int ROWS_BLOCK_MIN_L = 2048;  // shorter rows use the wave / thread kernels

static const RedCfg* red_cfgs(int fam, int& n) {
  switch (fam) {
    case VKH_RFAM_ALL: n = (int)(sizeof(ALL_CFGS) / sizeof(RedCfg)); return ALL_CFGS;
    case VKH_RFAM_ROWS: n = (int)(sizeof(ROWS_CFGS) / sizeof(RedCfg)); return ROWS_CFGS;
    case VKH_RFAM_COLS: n = (int)(sizeof(COLS_CFGS) / sizeof(RedCfg)); return COLS_CFGS;
  }
  n = 1;  // a single kernel choice
  return nullptr;
}

// family of a reduction: kind 0 = full (n = prev * K * post), 1 = axis
static int red_family(int kind, uint64_t K, uint64_t post) {
  if (kind == 0) return VKH_RFAM_ALL;
  if (post > 1) return VKH_RFAM_COLS;
  return K >= ROWS_BLOCK_MIN_L ? VKH_RFAM_ROWS : VKH_RFAM_NONE;
}

static std::array<uint64_t, 4> red_key(int fam, uint64_t prev, uint64_t K, uint64_t post) {
  if (fam == VKH_RFAM_ALL) return {(uint64_t)fam, 1, prev * K * post, 1};
  return {(uint64_t)fam, prev, K, post};
}

// configuration of a call: the explicit index when >= 0, else the one chosen
// for the shape on this context, else the default
static RedCfg red_cfg(vkh_gpu* g, int fam, uint64_t prev, uint64_t K, uint64_t post, int index) {
  int n = 0;
  const RedCfg* c = red_cfgs(fam, n);
  if (!c) return RedCfg{4, 0};
  if (index < 0) {
    auto it = g->red_cfg.find(red_key(fam, prev, K, post));
    index = it == g->red_cfg.end() ? 0 : it->second;
  }
  return c[index];
}

// blocks of a full reduction of n elements: one FP64 partial each
static unsigned all_blocks(vkh_gpu* g, uint64_t n, const RedCfg& c) {
  return stream_grid(g, ceil_div(n, 4 * 4), ALL_BLOCK, (unsigned)c.par);
}

// partial pass into part[0, nb), then the final combine into out[0]
template <int OP>
static void launch_all(vkh_gpu* g, const float* a, uint64_t n, float* out, int mean, const RedCfg& c,
                       unsigned nb, double* part) {
  if (c.unroll == 8)
    hipLaunchKernelGGL((k_all_partial<OP, 8>), dim3(nb), dim3(ALL_BLOCK), 0, g->stream, a, n, part);
  else
    hipLaunchKernelGGL((k_all_partial<OP, 4>), dim3(nb), dim3(ALL_BLOCK), 0, g->stream, a, n, part);
  check_launch("reduce_all partial");
  hipLaunchKernelGGL(k_all_final<OP>, dim3(1), dim3(ALL_BLOCK), 0, g->stream, part, (uint64_t)nb, out, n, mean);
  check_launch("reduce_all final");
}

template <int OP>
static void reduce_all(vkh_gpu* g, const float* a, uint64_t n, float* out, int mean, int index = -1) {
  const RedCfg c = red_cfg(g, VKH_RFAM_ALL, 1, n, 1, index);
  const unsigned nb = all_blocks(g, n, c);
  Workspace ws(g, sizeof(double) * nb);
  launch_all<OP>(g, a, n, out, mean, c, nb, (double*)ws.get());
}

// ---- allocation-free full reduction (vkh_reduce_all_noalloc) ------------------
// The FP64 block partials go to this file-scope device array instead of a
// workspace.  Each device holds its own copy at its own address, looked up
// once per device with hipGetSymbolAddress; host code passes that address,
// never the array's name (on the host the name is a stand-in, not device
// memory).  Every block writes its own partial before the final kernel reads
// it, so no call depends on what an earlier call left there.
constexpr unsigned NOALLOC_PARTIALS = 8192;  // 64 KiB: 256 CUs x 16 blocks per CU, twice over
static __device__ __align__(256) double scratch[NOALLOC_PARTIALS];

constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[MAX_DEVICES] = {};

// address of this device's scratch; the device must be current
static double* scratch_address(vkh_gpu* g) {
  VKH_REQUIRE(g->device >= 0 && g->device < MAX_DEVICES, "reduce_all_noalloc: device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  double*& p = g_scratch_ptr[g->device];
  if (!p) {
    void* addr = nullptr;
    VKH_HIP(hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch)));
    VKH_REQUIRE(addr != nullptr, "reduce_all_noalloc: no device address for the scratch array");
    p = (double*)addr;
  }
  return p;
}

template <int OP>
static void reduce_all_noalloc(vkh_gpu* g, const float* a, uint64_t n, float* out, int mean) {
  const RedCfg c = red_cfg(g, VKH_RFAM_ALL, 1, n, 1, -1);
  const unsigned nb = all_blocks(g, n, c);
  if (nb > NOALLOC_PARTIALS)
    fail(VKH_INVALID, "reduce_all_noalloc: " + std::to_string(nb) + " block partials exceed the " +
                          std::to_string(NOALLOC_PARTIALS) + " the scratch array holds");
  launch_all<OP>(g, a, n, out, mean, c, nb, scratch_address(g));
}

// Rows per chunk (and the chunk count) of the strided-axis reduction of
// [prev, K, post] in configuration cfg.  Split k when there are fewer than
// cfg.par threads of columns per CU and the axis is long enough to give every
// chunk a reasonable amount of work (default 512: on 8192 x 16384, 32 chunks
// of 256 rows reach 6.4 TB/s, 64 chunks 6.1).  The chunk count fixes the
// summation order of sums and products, so it is part of the shape's
// configuration.
static uint64_t cols_chunk(vkh_gpu* g, const RedCfg& cfg, uint64_t prev, uint64_t K, uint64_t post, bool vec,
                           uint64_t& nchunk_out) {
  const unsigned cus = g->num_cus > 0 ? g->num_cus : 256;
  const uint64_t threads = prev * (post / (vec ? 4 : 1));
  const uint64_t want = (uint64_t)cus * (uint64_t)cfg.par;
  uint64_t nchunk = 1;
  if (threads < want && K >= 64) {
    nchunk = std::min<uint64_t>(ceil_div(want, threads), K / 32);
    if (nchunk > 1024) nchunk = 1024;
    if (nchunk < 1) nchunk = 1;
  }
  const uint64_t kchunk = ceil_div(K, nchunk);
  nchunk_out = ceil_div(K, kchunk);
  return kchunk;
}

template <int OP>
static void reduce_axis(vkh_gpu* g, const float* a, float* out, uint64_t prev, uint64_t K,
                        uint64_t post, int rebroadcast, float mul, float dv, int index = -1) {
  const unsigned cus = g->num_cus > 0 ? g->num_cus : 256;
  const RedCfg cfg = red_cfg(g, red_family(1, K, post), prev, K, post, index);
  if (post == 1) {
    const uint64_t rows = prev, L = K;
    const bool vec = (L % 4 == 0) && (((uintptr_t)a & 15u) == 0) && (((uintptr_t)out & 15u) == 0);
    if (L >= ROWS_BLOCK_MIN_L) {
      unsigned grid = (unsigned)std::min<uint64_t>(rows, (uint64_t)cus * cfg.par);
      if (vec && cfg.unroll == 8)
        hipLaunchKernelGGL((k_rows_block<OP, true, 8>), dim3(grid), dim3(ROW_BLOCK), 0, g->stream, a, out, rows, L, rebroadcast, mul, dv);
      else if (vec)
        hipLaunchKernelGGL((k_rows_block<OP, true, 4>), dim3(grid), dim3(ROW_BLOCK), 0, g->stream, a, out, rows, L, rebroadcast, mul, dv);
      else
        hipLaunchKernelGGL((k_rows_block<OP, false, 4>), dim3(grid), dim3(ROW_BLOCK), 0, g->stream, a, out, rows, L, rebroadcast, mul, dv);
    } else if (L >= 64) {
      unsigned grid = (unsigned)std::min<uint64_t>(ceil_div(rows, 4), (uint64_t)cus * 16);
      hipLaunchKernelGGL(k_rows_wave<OP>, dim3(grid), dim3(256), 0, g->stream, a, out, rows, L, rebroadcast, mul, dv);
    } else {
      unsigned grid = (unsigned)std::min<uint64_t>(ceil_div(rows, 256), (uint64_t)cus * 16);
      hipLaunchKernelGGL(k_rows_thread<OP>, dim3(grid), dim3(256), 0, g->stream, a, out, rows, L, rebroadcast, mul, dv);
    }
    check_launch("reduce rows");
    return;
  }
  const bool vec = (post % 4 == 0) && (((uintptr_t)a & 15u) == 0) && (((uintptr_t)out & 15u) == 0);
  const int VW = vec ? 4 : 1;
  const uint64_t threads = prev * (post / VW);
  uint64_t nchunk = 0;
  const uint64_t kchunk = cols_chunk(g, cfg, prev, K, post, vec, nchunk);
  const uint64_t m = prev * post;
  Workspace ws(g, (nchunk > 1 || rebroadcast) ? sizeof(float) * nchunk * m : 1);
  float* dst = (nchunk > 1 || rebroadcast) ? (float*)ws.get() : out;
  if (vec && !(nchunk > 1 || rebroadcast)) { /* out alignment already checked */ }
  unsigned gx = (unsigned)std::min<uint64_t>(ceil_div(threads, 256), (uint64_t)cus * 16);
  if (gx < 1) gx = 1;
  dim3 grid(gx, (unsigned)nchunk);
  const int final_pass = (nchunk == 1) ? 1 : 0;
  if (VW == 4 && cfg.unroll == 8)
    hipLaunchKernelGGL((k_cols<OP, 4, 8>), grid, dim3(256), 0, g->stream, a, dst, prev, K, post, kchunk, final_pass, mul, dv);
  else if (VW == 4)
    hipLaunchKernelGGL((k_cols<OP, 4, 4>), grid, dim3(256), 0, g->stream, a, dst, prev, K, post, kchunk, final_pass, mul, dv);
  else
    hipLaunchKernelGGL((k_cols<OP, 1, 4>), grid, dim3(256), 0, g->stream, a, dst, prev, K, post, kchunk, final_pass, mul, dv);
  check_launch("reduce cols");
  float* red = dst;
  if (nchunk > 1) {
    // combine into the first chunk's slot when rebroadcasting, else into out
    float* target = rebroadcast ? dst : out;
    if (rebroadcast) {
      // the combine reads every chunk, so write to a separate buffer
      Workspace ws2(g, sizeof(float) * m);
      float* t2 = (float*)ws2.get();
      unsigned gb = (unsigned)std::min<uint64_t>(ceil_div(m, 256), (uint64_t)cus * 16);
      hipLaunchKernelGGL(k_combine_chunks<OP>, dim3(gb), dim3(256), 0, g->stream, dst, t2, m, nchunk, mul, dv);
      check_launch("reduce combine");
      unsigned gr = stream_grid(g, prev * K * post, 256, 8);
      hipLaunchKernelGGL(k_rebroadcast_cols, dim3(gr), dim3(256), 0, g->stream, t2, out, prev, K, post);
      check_launch("reduce rebroadcast");
      return;
    }
    unsigned gb = (unsigned)std::min<uint64_t>(ceil_div(m, 256), (uint64_t)cus * 16);
    hipLaunchKernelGGL(k_combine_chunks<OP>, dim3(gb), dim3(256), 0, g->stream, dst, target, m, nchunk, mul, dv);
    check_launch("reduce combine");
    return;
  }
  if (rebroadcast) {
    unsigned gr = stream_grid(g, prev * K * post, 256, 8);
    hipLaunchKernelGGL(k_rebroadcast_cols, dim3(gr), dim3(256), 0, g->stream, red, out, prev, K, post);
    check_launch("reduce rebroadcast");
  }
}

}  // namespace vkh

using namespace vkh;

extern "C" {

int vkh_reduce_all(vkh_gpu* g, int op, const float* a, uint64_t n, float* out, int mean) {
  return guarded([&] {
    use(g);
    VKH_REQUIRE(out, "reduce_all: null output");
    VKH_REQUIRE(n == 0 || a, "reduce_all: null input");
    if (n == 0) {
      VKH_REQUIRE(op == VKH_RSUM || op == VKH_RPROD,
                  "zero-size array to reduction operation which has no identity");
    }
    switch (op) {
      case VKH_RSUM: reduce_all<VKH_RSUM>(g, a, n, out, mean); break;
      case VKH_RPROD: reduce_all<VKH_RPROD>(g, a, n, out, mean); break;
      case VKH_RMAX: reduce_all<VKH_RMAX>(g, a, n, out, mean); break;
      case VKH_RMIN: reduce_all<VKH_RMIN>(g, a, n, out, mean); break;
      default: fail(VKH_INVALID, "unknown reduction op");
    }
  });
}

}  // extern "C"

namespace vkh {
// chunk length of the column sums of a [rows, cols] operand (vkh_bias_grad's
// reduction), for kernels that must sum in the same order; vec: the operand
// and output allow 16-byte accesses
uint64_t bias_grad_chunk(vkh_gpu* g, uint64_t rows, uint64_t cols, bool vec) {
  const RedCfg cfg = red_cfg(g, red_family(1, rows, cols), 1, rows, cols, -1);
  uint64_t nchunk = 0;
  return cols_chunk(g, cfg, 1, rows, cols, vec, nchunk);
}

// Chunk partials part[c][m] summed per output in chunk order: FROM_ZERO 0
// starts from part[0] (k_combine_chunks' order for a sum, no scaling), 1 from
// +0 (the column walk's combine).  For few outputs and many chunks: a block
// takes 64 outputs, its four waves load batches of up to 128 chunks (256-byte
// rows, every load of a thread issued before the first wait) into LDS, and
// lane j of the first wave adds output j's batch in order.  The same sums as
// the one-thread-per-output combines, with all their loads in flight at once.
constexpr int CC_OUT = 64, CC_BATCH = 128, CC_PER = CC_BATCH / 4;
template <int FROM_ZERO>
__global__ void __launch_bounds__(256)
k_combine_cols(const float* __restrict__ part, float* __restrict__ out, uint64_t m, uint64_t nchunk) {
  __shared__ float tile[CC_BATCH][CC_OUT + 1];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t j = (uint64_t)blockIdx.x * CC_OUT + lane;
  float acc = 0.f;
  for (uint64_t c0 = 0; c0 < nchunk; c0 += CC_BATCH) {
    float r[CC_PER];
#pragma unroll
    for (int q = 0; q < CC_PER; ++q) {
      const uint64_t c = c0 + w + 4 * q;
      r[q] = (c < nchunk && j < m) ? part[c * m + j] : 0.f;
    }
#pragma unroll
    for (int q = 0; q < CC_PER; ++q) tile[w + 4 * q][lane] = r[q];
    __syncthreads();
    if (w == 0) {
      const int nb = (int)min<uint64_t>(CC_BATCH, nchunk - c0);
      for (int c = 0; c < nb; ++c) {
        const float v = tile[c][lane];
        acc = (!FROM_ZERO && c0 == 0 && c == 0) ? v : acc + v;
      }
    }
    __syncthreads();
  }
  if (w == 0 && j < m) out[j] = acc;
}

static void combine_cols(vkh_gpu* g, const float* part, float* out, uint64_t m, uint64_t nchunk, bool from_zero) {
  const unsigned grid = (unsigned)ceil_div(m, (uint64_t)CC_OUT);
  if (from_zero) hipLaunchKernelGGL(k_combine_cols<1>, dim3(grid), dim3(256), 0, g->stream, part, out, m, nchunk);
  else hipLaunchKernelGGL(k_combine_cols<0>, dim3(grid), dim3(256), 0, g->stream, part, out, m, nchunk);
  check_launch("combine chunks");
}

// combine of chunk partials [nchunk, m] in chunk order (k_combine_chunks' sums)
void sum_chunks(vkh_gpu* g, const float* part, float* out, uint64_t m, uint64_t nchunk) {
  combine_cols(g, part, out, m, nchunk, false);
}

// the same from +0 (the skinny column walk's combine, k_skinny_combine with
// alpha 1, beta 0 and no epilogue)
void sum_chunks_from_zero(vkh_gpu* g, const float* part, float* out, uint64_t m, uint64_t nchunk) {
  combine_cols(g, part, out, m, nchunk, true);
}
}  // namespace vkh

extern "C" {

int vkh_reduce_all_noalloc(vkh_gpu* g, int op, const float* a, uint64_t n, float* out, int mean) {
  return guarded([&] {
    use(g);
    VKH_REQUIRE(out, "reduce_all_noalloc: null output");
    VKH_REQUIRE(n == 0 || a, "reduce_all_noalloc: null input");
    if (n == 0) {
      VKH_REQUIRE(op == VKH_RSUM || op == VKH_RPROD,
                  "zero-size array to reduction operation which has no identity");
    }
    switch (op) {
      case VKH_RSUM: reduce_all_noalloc<VKH_RSUM>(g, a, n, out, mean); break;
      case VKH_RPROD: reduce_all_noalloc<VKH_RPROD>(g, a, n, out, mean); break;
      case VKH_RMAX: reduce_all_noalloc<VKH_RMAX>(g, a, n, out, mean); break;
      case VKH_RMIN: reduce_all_noalloc<VKH_RMIN>(g, a, n, out, mean); break;
      default: fail(VKH_INVALID, "unknown reduction op");
    }
  });
}

int vkh_reduce_axis(vkh_gpu* g, int op, const float* a, float* out, uint64_t prev, uint64_t axis,
                    uint64_t post, int rebroadcast, float mul, float dv) {
  return guarded([&] {
    use(g);
    VKH_REQUIRE(a && out, "reduce_axis: null pointer");
    if (prev == 0 || post == 0) return;
    if (axis == 0) {
      VKH_REQUIRE(op == VKH_RSUM || op == VKH_RPROD,
                  "zero-size array to reduction operation which has no identity");
      // empty axis: identity, scaled like any result
      float ident = op == VKH_RSUM ? 0.0f : 1.0f;
      if (mul != 1.0f) ident *= mul;
      if (dv != 1.0f) ident /= dv;
      if (!rebroadcast) {
        int r = vkh_fill_f32(g, out, ident, prev * post);
        if (r != VKH_OK) fail(r, vkh_last_error());
      }
      return;
    }
    switch (op) {
      case VKH_RSUM: reduce_axis<VKH_RSUM>(g, a, out, prev, axis, post, rebroadcast, mul, dv); break;
      case VKH_RPROD: reduce_axis<VKH_RPROD>(g, a, out, prev, axis, post, rebroadcast, mul, dv); break;
      case VKH_RMAX: reduce_axis<VKH_RMAX>(g, a, out, prev, axis, post, rebroadcast, mul, dv); break;
      case VKH_RMIN: reduce_axis<VKH_RMIN>(g, a, out, prev, axis, post, rebroadcast, mul, dv); break;
      default: fail(VKH_INVALID, "unknown reduction op");
    }
  });
}

int vkh_reduce_family(int kind, uint64_t prev, uint64_t axis, uint64_t post, int* family) {
  return guarded([&] {
    VKH_REQUIRE(family, "null out");
    VKH_REQUIRE(kind == 0 || kind == 1, "reduction kind must be 0 (full) or 1 (axis)");
    (void)prev;
    *family = red_family(kind, axis, post);
  });
}

int vkh_reduce_configs(int family, int* count) {
  return guarded([&] {
    VKH_REQUIRE(count, "null out");
    VKH_REQUIRE(family >= VKH_RFAM_NONE && family <= VKH_RFAM_COLS, "unknown reduction family");
    red_cfgs(family, *count);
  });
}

int vkh_reduce_config_info(int family, int index, int* unroll, int* par) {
  return guarded([&] {
    VKH_REQUIRE(unroll && par, "null out");
    VKH_REQUIRE(family >= VKH_RFAM_NONE && family <= VKH_RFAM_COLS, "unknown reduction family");
    int n = 0;
    const RedCfg* c = red_cfgs(family, n);
    VKH_REQUIRE(index >= 0 && index < n, "configuration index out of range");
    *unroll = c ? c[index].unroll : 0;
    *par = c ? c[index].par : 0;
  });
}

int vkh_reduce_get_config(vkh_gpu* g, int kind, uint64_t prev, uint64_t axis, uint64_t post,
                          int* index, int* tuned) {
  return guarded([&] {
    VKH_REQUIRE(g && index, "null argument");
    VKH_REQUIRE(kind == 0 || kind == 1, "reduction kind must be 0 (full) or 1 (axis)");
    const int fam = red_family(kind, axis, post);
    auto it = g->red_cfg.find(red_key(fam, prev, axis, post));
    const bool have = it != g->red_cfg.end();
    *index = have ? it->second : 0;
    if (tuned) *tuned = have ? 1 : 0;
  });
}

int vkh_reduce_set_config(vkh_gpu* g, int kind, uint64_t prev, uint64_t axis, uint64_t post,
                          int index) {
  return guarded([&] {
    VKH_REQUIRE(g, "null GPU handle");
    VKH_REQUIRE(kind == 0 || kind == 1, "reduction kind must be 0 (full) or 1 (axis)");
    const int fam = red_family(kind, axis, post);
    const auto key = red_key(fam, prev, axis, post);
    if (index < 0) { g->red_cfg.erase(key); return; }
    int n = 0;
    red_cfgs(fam, n);
    VKH_REQUIRE(index < n, "configuration index out of range");
    if (fam == VKH_RFAM_NONE) return;  // a single kernel choice: nothing to keep
    g->red_cfg[key] = index;
  });
}

int vkh_reduce_autotune(vkh_gpu* g, int kind, uint64_t prev, uint64_t axis, uint64_t post,
                        int reps, int same_order, int* best, float* ms_per_config) {
  return guarded([&] {
    use(g);
    VKH_REQUIRE(kind == 0 || kind == 1, "reduction kind must be 0 (full) or 1 (axis)");
    VKH_REQUIRE(!g->capturing, "reduce_autotune: not allowed during graph capture");
    VKH_REQUIRE(prev > 0 && axis > 0 && post > 0, "reduce_autotune: empty reduction");
    VKH_REQUIRE(g->alloc, "reduce_autotune needs a context with its own allocator");
    if (reps < 1) reps = 3;
    const int fam = red_family(kind, axis, post);
    int n = 0;
    const RedCfg* cfgs = red_cfgs(fam, n);
    if (fam == VKH_RFAM_NONE) {  // a single kernel choice
      if (ms_per_config) ms_per_config[0] = -1.0f;
      if (best) *best = 0;
      return;
    }
    // scratch input and output of the reduction's own shape
    const uint64_t total = prev * axis * post;
    const uint64_t nout = kind == 0 ? 1 : prev * post;
    Workspace win(g, sizeof(float) * total), wout(g, sizeof(float) * nout);
    float* A = (float*)win.get();
    float* O = (float*)wout.get();
    hipLaunchKernelGGL(k_fill_noise, dim3(full_grid(total, 256)), dim3(256), 0, g->stream, A, total, 5u);
    check_launch("reduce_autotune fill");
    auto run = [&](int i) {
      if (kind == 0) reduce_all<VKH_RSUM>(g, A, total, O, 0, i);
      else reduce_axis<VKH_RSUM>(g, A, O, prev, axis, post, 0, 1.0f, 1.0f, i);
    };
    hipEvent_t e0, e1;
    VKH_HIP(hipEventCreate(&e0));
    VKH_HIP(hipEventCreate(&e1));
    int bi = -1;
    float bt = 0.0f;
    // same_order: only configurations whose summation order is the
    // default's (every ROWS configuration; ALL and COLS with the default
    // par), so the choice cannot change a result
    auto candidate = [&](int i) { return !same_order || fam == VKH_RFAM_ROWS || cfgs[i].par == cfgs[0].par; };
    try {
      for (int i = 0; i < n; ++i) {
        if (!candidate(i)) {
          if (ms_per_config) ms_per_config[i] = -1.0f;
          continue;
        }
        run(i);  // warm-up
        VKH_HIP(hipEventRecord(e0, g->stream));
        for (int r = 0; r < reps; ++r) run(i);
        VKH_HIP(hipEventRecord(e1, g->stream));
        VKH_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        VKH_HIP(hipEventElapsedTime(&ms, e0, e1));
        ms /= (float)reps;
        if (ms_per_config) ms_per_config[i] = ms;
        if (bi < 0 || ms < bt) { bi = i; bt = ms; }
      }
    } catch (...) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    g->red_cfg[red_key(fam, prev, axis, post)] = bi;
    if (best) *best = bi;
  });
}

}  // extern "C"
