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
template <int OP, bool VEC>
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
      // four loads in flight, combined in the same order as the plain loop
      for (; i + 3 * ROW_BLOCK < L4; i += 4 * ROW_BLOCK) {
        const float4 v0 = ld_stream(row4 + i), v1 = ld_stream(row4 + i + ROW_BLOCK);
        const float4 v2 = ld_stream(row4 + i + 2 * ROW_BLOCK), v3 = ld_stream(row4 + i + 3 * ROW_BLOCK);
        acc = R::f(acc, R::f(R::f(v0.x, v0.y), R::f(v0.z, v0.w)));
        acc = R::f(acc, R::f(R::f(v1.x, v1.y), R::f(v1.z, v1.w)));
        acc = R::f(acc, R::f(R::f(v2.x, v2.y), R::f(v2.z, v2.w)));
        acc = R::f(acc, R::f(R::f(v3.x, v3.y), R::f(v3.z, v3.w)));
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
// post + j] where c is the chunk index (blockIdx.y).
template <int OP, int VW>
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
      // four rows in flight, still combined in row order
      for (; k + 3 < k1; k += 4) {
        const float4 v0 = ld_stream(reinterpret_cast<const float4*>(base + k * post));
        const float4 v1 = ld_stream(reinterpret_cast<const float4*>(base + (k + 1) * post));
        const float4 v2 = ld_stream(reinterpret_cast<const float4*>(base + (k + 2) * post));
        const float4 v3 = ld_stream(reinterpret_cast<const float4*>(base + (k + 3) * post));
        comb(acc, v0);
        comb(acc, v1);
        comb(acc, v2);
        comb(acc, v3);
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

template <int OP>
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

template <int OP>
static void reduce_all(vkh_gpu* g, const float* a, uint64_t n, float* out, int mean) {
  unsigned nb = stream_grid(g, ceil_div(n, 4 * 4), ALL_BLOCK, 4);
  Workspace ws(g, sizeof(double) * nb);
  double* part = (double*)ws.get();
  hipLaunchKernelGGL(k_all_partial<OP>, dim3(nb), dim3(ALL_BLOCK), 0, g->stream, a, n, part);
  check_launch("reduce_all partial");
  hipLaunchKernelGGL(k_all_final<OP>, dim3(1), dim3(ALL_BLOCK), 0, g->stream, part, (uint64_t)nb, out, n, mean);
  check_launch("reduce_all final");
}

template <int OP>
static void reduce_axis(vkh_gpu* g, const float* a, float* out, uint64_t prev, uint64_t K,
                        uint64_t post, int rebroadcast, float mul, float dv) {
  const unsigned cus = g->num_cus > 0 ? g->num_cus : 256;
  if (post == 1) {
    const uint64_t rows = prev, L = K;
    const bool vec = (L % 4 == 0) && (((uintptr_t)a & 15u) == 0) && (((uintptr_t)out & 15u) == 0);
    if (L >= 2048) {
      unsigned grid = (unsigned)std::min<uint64_t>(rows, (uint64_t)cus * 16);
      if (vec)
        hipLaunchKernelGGL((k_rows_block<OP, true>), dim3(grid), dim3(ROW_BLOCK), 0, g->stream, a, out, rows, L, rebroadcast, mul, dv);
      else
        hipLaunchKernelGGL((k_rows_block<OP, false>), dim3(grid), dim3(ROW_BLOCK), 0, g->stream, a, out, rows, L, rebroadcast, mul, dv);
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
  // Split k when there are fewer than ~2 waves of columns per CU and the axis
  // is long enough to give every chunk a reasonable amount of work (measured
  // on 8192 x 16384: 32 chunks of 256 rows reach 6.4 TB/s, 64 chunks 6.1).
  const uint64_t want = (uint64_t)cus * 512;
  uint64_t nchunk = 1;
  if (threads < want && K >= 64) {
    nchunk = std::min<uint64_t>(ceil_div(want, threads), K / 32);
    if (nchunk > 1024) nchunk = 1024;
    if (nchunk < 1) nchunk = 1;
  }
  const uint64_t kchunk = ceil_div(K, nchunk);
  nchunk = ceil_div(K, kchunk);
  const uint64_t m = prev * post;
  Workspace ws(g, (nchunk > 1 || rebroadcast) ? sizeof(float) * nchunk * m : 1);
  float* dst = (nchunk > 1 || rebroadcast) ? (float*)ws.get() : out;
  if (vec && !(nchunk > 1 || rebroadcast)) { /* out alignment already checked */ }
  unsigned gx = (unsigned)std::min<uint64_t>(ceil_div(threads, 256), (uint64_t)cus * 16);
  if (gx < 1) gx = 1;
  dim3 grid(gx, (unsigned)nchunk);
  const int final_pass = (nchunk == 1) ? 1 : 0;
  if (VW == 4)
    hipLaunchKernelGGL((k_cols<OP, 4>), grid, dim3(256), 0, g->stream, a, dst, prev, K, post, kchunk, final_pass, mul, dv);
  else
    hipLaunchKernelGGL((k_cols<OP, 1>), grid, dim3(256), 0, g->stream, a, dst, prev, K, post, kchunk, final_pass, mul, dv);
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

}  // extern "C"
