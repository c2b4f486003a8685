// Elementwise kernels: same-shape, scalar, broadcast binary ops (F1-F6),
// unary math (F7/F8), clamp (F9) and broadcast copies (F10).
//
// The reference runs one thread per element with a per-dimension div/mod
// decode of a uint32 shape table for every broadcast op
// (vulkpy/shader/add_broadcast.comp:25-44).  Here broadcasting is planned on
// the host: size-1 dims are dropped and adjacent dims with compatible strides
// are merged, so the common cases become a 2-D [rows, cols] traversal where
// each operand is FULL (same layout as the output), ROW (indexed by column),
// COL (indexed by row) or SCALAR.  Those run with 128-bit loads/stores and a
// grid-stride loop; anything else takes a general strided N-d kernel.
#include <hip/hip_ext.h>

#include "bmath.h"
#include "common.h"
#include "fastdiv.h"

namespace vkh {

// Operand kinds of the 2-D traversal of out[rows, cols]:
//   FULL   p[row * cols + col]   (same layout as the output)
//   ROW    p[col]                (broadcast along rows, shape [1, cols])
//   COL    p[row]                (broadcast along columns, shape [rows, 1])
//   FIRST  p[0]                  (one element broadcast everywhere)
//   SCALAR s                     (host scalar)
enum Kind : int { K_FULL = 0, K_ROW = 1, K_COL = 2, K_FIRST = 3, K_SCALAR = 4 };

struct Opnd {
  const float* p;
  float s;
  int kind;
};

// ---- functors ------------------------------------------------------------------
template <int OP>
struct BinF {
  static constexpr int nin = 2;
  __device__ __forceinline__ float operator()(float a, float b, float) const { return binop<OP>(a, b); }
};

// a ** s for the scalar exponents whose power is a single correctly rounded
// FP32 operation, as NumPy's scalar power does: the FP64 pow rounded once
// (pow's special cases included; exact ties round to even, which the general
// exp(b log a) evaluation can only approach) at the cost of that operation.
// E: 1 -> a, 2 -> a * a, -1 -> 1 / a, 0 -> 1, 5 -> sqrt(a) (pow(-0, 0.5) = +0
// and pow(-inf, 0.5) = +inf, where sqrt gives -0 and NaN)
template <int E>
struct PowScalarF {
  static constexpr int nin = 2;
  __device__ __forceinline__ float operator()(float a, float, float) const {
    if constexpr (E == 1) return a;
    else if constexpr (E == 2) return a * a;
    else if constexpr (E == -1) return 1.0f / a;
    else if constexpr (E == 0) return 1.0f;
    else {
      if (a == 0.0f) return 0.0f;
      if (a == -INFINITY) return INFINITY;
      return sqrtf(a);
    }
  }
};

struct ClampF {
  static constexpr int nin = 3;
  // GLSL clamp(x, lo, hi) == min(max(x, lo), hi) == numpy.clip
  __device__ __forceinline__ float operator()(float x, float lo, float hi) const {
    return np_min(np_max(x, lo), hi);
  }
};

struct CopyF {
  static constexpr int nin = 1;
  __device__ __forceinline__ float operator()(float a, float, float) const { return a; }
};

__device__ __forceinline__ float4 load4(const Opnd& o, uint64_t t, uint64_t row, uint64_t col4) {
  switch (o.kind) {
    case K_FULL: return reinterpret_cast<const float4*>(o.p)[t];
    case K_ROW: return reinterpret_cast<const float4*>(o.p)[col4];
    case K_COL: { float x = o.p[row]; return make_float4(x, x, x, x); }
    case K_FIRST: { float x = o.p[0]; return make_float4(x, x, x, x); }
    default: return make_float4(o.s, o.s, o.s, o.s);
  }
}

__device__ __forceinline__ float load1(const Opnd& o, uint64_t t, uint64_t row, uint64_t col) {
  switch (o.kind) {
    case K_FULL: return o.p[t];
    case K_ROW: return o.p[col];
    case K_COL: return o.p[row];
    case K_FIRST: return o.p[0];
    default: return o.s;
  }
}

template <class F>
__device__ __forceinline__ float4 apply4(const F& f, float4 a, float4 b, float4 c) {
  return make_float4(f(a.x, b.x, c.x), f(a.y, b.y, c.y), f(a.z, b.z, c.z), f(a.w, b.w, c.w));
}

constexpr int EW_BLOCK = 256;

// Vectorised 2-D kernel: the first total4 float4 units of out, where out is
// [rows, cols] with cols % 4 == 0 (or rows == 1) and cols4 = cols / 4.  One
// unit per thread over a full grid (the loop only runs again for arrays
// beyond 2^24 blocks); full-size operands and the output stream with
// nontemporal accesses, broadcast operands use ordinary (cached) loads.  Row
// and column of unit t come from one fast division (total4 < 2^31).
template <class F, bool NEED_ROWCOL>
__global__ void __launch_bounds__(EW_BLOCK)
k_ew2d_vec(F f, Opnd a, Opnd b, Opnd c, float* out, uint32_t total4, FastDiv cols4) {
  const uint32_t stride = gridDim.x * EW_BLOCK;
  for (uint32_t t = blockIdx.x * EW_BLOCK + threadIdx.x; t < total4; t += stride) {
    uint32_t row = 0, col4 = t;
    if constexpr (NEED_ROWCOL) { row = cols4.div(t); col4 = t - row * cols4.d; }
    const float4 va = a.kind == K_FULL ? ld_stream(reinterpret_cast<const float4*>(a.p) + t) : load4(a, t, row, col4);
    float4 vb = va, vc = va;
    if (F::nin > 1) vb = b.kind == K_FULL ? ld_stream(reinterpret_cast<const float4*>(b.p) + t) : load4(b, t, row, col4);
    if (F::nin > 2) vc = c.kind == K_FULL ? ld_stream(reinterpret_cast<const float4*>(c.p) + t) : load4(c, t, row, col4);
    st_stream(reinterpret_cast<float4*>(out) + t, apply4(f, va, vb, vc));
  }
}

// The same traversal with the operand kinds fixed at compile time.  With the
// kinds known, the loads of all operands of a unit issue back to back; the
// run-time switch of k_ew2d_vec splits the code into branches whose merge
// points wait for each load before the next is issued.
template <int KIND>
__device__ __forceinline__ float4 load4k(const Opnd& o, uint32_t t, uint32_t row, uint32_t col4) {
  if constexpr (KIND == K_FULL) return ld_stream(reinterpret_cast<const float4*>(o.p) + t);
  else if constexpr (KIND == K_ROW) return reinterpret_cast<const float4*>(o.p)[col4];
  else if constexpr (KIND == K_COL) { const float x = o.p[row]; return make_float4(x, x, x, x); }
  else if constexpr (KIND == K_FIRST) { const float x = o.p[0]; return make_float4(x, x, x, x); }
  else return make_float4(o.s, o.s, o.s, o.s);
}

template <class F, int KA, int KB, int KC>
__global__ void __launch_bounds__(EW_BLOCK)
k_ew2d_kind(F f, Opnd a, Opnd b, Opnd c, float* out, uint32_t total4, FastDiv cols4) {
  constexpr bool RC = KA == K_ROW || KA == K_COL || KB == K_ROW || KB == K_COL || KC == K_ROW || KC == K_COL;
  const uint32_t stride = gridDim.x * EW_BLOCK;
  for (uint32_t t = blockIdx.x * EW_BLOCK + threadIdx.x; t < total4; t += stride) {
    uint32_t row = 0, col4 = t;
    if constexpr (RC) { row = cols4.div(t); col4 = t - row * cols4.d; }
    const float4 va = load4k<KA>(a, t, row, col4);
    const float4 vb = F::nin > 1 ? load4k<KB>(b, t, row, col4) : va;
    const float4 vc = F::nin > 2 ? load4k<KC>(c, t, row, col4) : va;
    st_stream(reinterpret_cast<float4*>(out) + t, apply4(f, va, vb, vc));
  }
}

// Every elementwise launch goes through here.  When vkh_timer_next_kernel has
// armed the context's timing events, the kernel is launched with
// hipExtLaunchKernelGGL, whose dispatch records them at the kernel's own
// start and end, and the events are disarmed; otherwise a plain launch.
template <typename... KArgs, typename... Args>
static void ew_launch(vkh_gpu* g, void (*kernel)(KArgs...), dim3 grid, Args... args) {
  if (g->ktime_start && !g->capturing) {
    hipEvent_t a = g->ktime_start, b = g->ktime_stop;
    g->ktime_start = g->ktime_stop = nullptr;
    hipExtLaunchKernelGGL(kernel, grid, dim3(EW_BLOCK), 0, g->stream, a, b, 0, args...);
  } else {
    hipLaunchKernelGGL(kernel, grid, dim3(EW_BLOCK), 0, g->stream, args...);
  }
}

// Launches k_ew2d_kind when the operand kinds are one of the streaming
// combinations (at least one full-size operand); returns false otherwise.
template <class F>
static bool launch_kind(dim3 grid, vkh_gpu* g, const F& f, const Opnd* op, float* out, uint32_t total4,
                        FastDiv fd) {
  const int ka = op[0].kind, kb = op[1].kind, kc = op[2].kind;
#define VKH_KIND(A, B, C)                                                                     \
  if (ka == A && kb == B && kc == C) {                                                        \
    ew_launch(g, k_ew2d_kind<F, A, B, C>, grid, f, op[0], op[1], op[2], out, total4, fd);     \
    return true;                                                                              \
  }
  if constexpr (F::nin == 1) {
    VKH_KIND(K_FULL, K_SCALAR, K_SCALAR)
  } else if constexpr (F::nin == 2) {
    VKH_KIND(K_FULL, K_FULL, K_SCALAR)
    VKH_KIND(K_FULL, K_COL, K_SCALAR)
    VKH_KIND(K_FULL, K_ROW, K_SCALAR)
    VKH_KIND(K_FULL, K_SCALAR, K_SCALAR)
    VKH_KIND(K_COL, K_FULL, K_SCALAR)
    VKH_KIND(K_ROW, K_FULL, K_SCALAR)
    VKH_KIND(K_SCALAR, K_FULL, K_SCALAR)
  } else {
    VKH_KIND(K_FULL, K_SCALAR, K_SCALAR)
    VKH_KIND(K_FULL, K_FULL, K_FULL)
  }
#undef VKH_KIND
  return false;
}

// Same traversal with 64-bit indices for 2^31 float4 units or more.
template <class F>
__global__ void __launch_bounds__(EW_BLOCK)
k_ew2d_vec64(F f, Opnd a, Opnd b, Opnd c, float* out, uint64_t total4, uint64_t cols4) {
  const uint64_t stride = (uint64_t)gridDim.x * EW_BLOCK;
  for (uint64_t t = (uint64_t)blockIdx.x * EW_BLOCK + threadIdx.x; t < total4; t += stride) {
    uint64_t row = t / cols4, col4 = t - row * cols4;
    const float4 va = a.kind == K_FULL ? ld_stream(reinterpret_cast<const float4*>(a.p) + t) : load4(a, t, row, col4);
    float4 vb = va, vc = va;
    if (F::nin > 1) vb = b.kind == K_FULL ? ld_stream(reinterpret_cast<const float4*>(b.p) + t) : load4(b, t, row, col4);
    if (F::nin > 2) vc = c.kind == K_FULL ? ld_stream(reinterpret_cast<const float4*>(c.p) + t) : load4(c, t, row, col4);
    st_stream(reinterpret_cast<float4*>(out) + t, apply4(f, va, vb, vc));
  }
}

// Scalar 2-D kernel over elements [begin, end) of out [rows, cols].
template <class F>
__global__ void __launch_bounds__(EW_BLOCK)
k_ew2d_scalar(F f, Opnd a, Opnd b, Opnd c, float* out, uint64_t begin, uint64_t end, uint64_t cols) {
  const uint64_t stride = (uint64_t)gridDim.x * EW_BLOCK;
  for (uint64_t t = begin + (uint64_t)blockIdx.x * EW_BLOCK + threadIdx.x; t < end; t += stride) {
    uint64_t row = t / cols, col = t - row * cols;
    float va = load1(a, t, row, col);
    float vb = F::nin > 1 ? load1(b, t, row, col) : va;
    float vc = F::nin > 2 ? load1(c, t, row, col) : va;
    out[t] = f(va, vb, vc);
  }
}

// General strided N-d kernel (after dimension collapsing).
constexpr int MAXD = 16;
struct NdDesc {
  int nd;
  uint64_t dims[MAXD];
  uint64_t st[3][MAXD];  // element strides per input (0 = broadcast)
  float s[3];            // scalar values for inputs with null pointers
  const float* p[3];
};

template <class F>
__global__ void __launch_bounds__(EW_BLOCK)
k_ewnd(F f, NdDesc d, float* out, uint64_t total) {
  const uint64_t stride = (uint64_t)gridDim.x * EW_BLOCK;
  for (uint64_t t = (uint64_t)blockIdx.x * EW_BLOCK + threadIdx.x; t < total; t += stride) {
    uint64_t rem = t, off0 = 0, off1 = 0, off2 = 0;
    for (int k = d.nd - 1; k >= 0; --k) {
      uint64_t q = rem / d.dims[k];
      uint64_t coord = rem - q * d.dims[k];
      rem = q;
      off0 += coord * d.st[0][k];
      off1 += coord * d.st[1][k];
      off2 += coord * d.st[2][k];
    }
    float va = d.p[0] ? d.p[0][off0] : d.s[0];
    float vb = F::nin > 1 ? (d.p[1] ? d.p[1][off1] : d.s[1]) : va;
    float vc = F::nin > 2 ? (d.p[2] ? d.p[2][off2] : d.s[2]) : va;
    out[t] = f(va, vb, vc);
  }
}

// ---- host-side broadcast planning -----------------------------------------------
struct InSpec {
  const float* p;           // null => scalar s
  float s;
  std::vector<uint64_t> shape;  // padded to ndim (ignored for scalars)
};

struct Plan {
  std::vector<uint64_t> dims;                // collapsed output dims (no 1s)
  std::vector<std::vector<uint64_t>> st;     // per input strides
  uint64_t total = 1;
};

static Plan make_plan(const std::vector<InSpec>& ins, const std::vector<uint64_t>& oshape) {
  const int nd = (int)oshape.size();
  Plan pl;
  for (uint64_t v : oshape) pl.total *= v;
  // element strides of every input in the output's index space (0 = broadcast)
  std::vector<std::vector<uint64_t>> st(ins.size(), std::vector<uint64_t>(nd, 0));
  for (size_t k = 0; k < ins.size(); ++k) {
    if (!ins[k].p) continue;
    VKH_REQUIRE((int)ins[k].shape.size() == nd, "operand rank mismatch");
    uint64_t s = 1;
    for (int d = nd - 1; d >= 0; --d) {
      uint64_t sd = ins[k].shape[d];
      if (sd == oshape[d]) st[k][d] = (sd == 1) ? 0 : s;
      else if (sd == 1) st[k][d] = 0;
      else fail(VKH_INVALID, "operands could not be broadcast together");
      s *= sd;
    }
  }
  // Drop size-1 output dims; merge a dim into the collapsed dim outside it
  // when every input has outer_stride == inner_stride * inner_size.
  pl.st.assign(ins.size(), {});
  for (int d = 0; d < nd; ++d) {
    if (oshape[d] == 1) continue;
    bool merge = !pl.dims.empty();
    for (size_t k = 0; merge && k < ins.size(); ++k)
      if (pl.st[k].back() != st[k][d] * oshape[d]) merge = false;
    if (merge) {
      pl.dims.back() *= oshape[d];
      for (size_t k = 0; k < ins.size(); ++k) pl.st[k].back() = st[k][d];
    } else {
      pl.dims.push_back(oshape[d]);
      for (size_t k = 0; k < ins.size(); ++k) pl.st[k].push_back(st[k][d]);
    }
  }
  return pl;
}

static bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

// Launch functor F over inputs (1..3) into out of shape oshape.
template <class F>
static void run_ew(vkh_gpu* g, const F& f, const std::vector<InSpec>& ins,
                   float* out, const std::vector<uint64_t>& oshape) {
  Plan pl = make_plan(ins, oshape);
  if (pl.total == 0) return;
  VKH_REQUIRE(out != nullptr, "null output pointer");
  VKH_REQUIRE(ins.size() >= 1 && ins.size() <= 3, "1 to 3 inputs");
  const size_t nin = ins.size();
  const int nd = (int)pl.dims.size();

  if (nd <= 2) {
    const uint64_t rows = nd == 2 ? pl.dims[0] : 1;
    const uint64_t cols = nd == 0 ? 1 : pl.dims[nd - 1];
    Opnd op[3] = {{nullptr, 0.f, K_SCALAR}, {nullptr, 0.f, K_SCALAR}, {nullptr, 0.f, K_SCALAR}};
    bool vec_ok = aligned16(out);
    bool need_rc = false;
    for (size_t k = 0; k < nin; ++k) {
      op[k].p = ins[k].p;
      op[k].s = ins[k].s;
      if (!ins[k].p) { op[k].kind = K_SCALAR; continue; }
      const auto& s = pl.st[k];
      if (nd == 0) op[k].kind = K_FIRST;
      else if (nd == 1) op[k].kind = s[0] == 1 ? K_FULL : K_FIRST;
      else if (s[0] == cols && s[1] == 1) op[k].kind = K_FULL;
      else if (s[0] == 0 && s[1] == 1) op[k].kind = K_ROW;
      else if (s[0] == 1 && s[1] == 0) op[k].kind = K_COL;
      else if (s[0] == 0 && s[1] == 0) op[k].kind = K_FIRST;
      else fail(VKH_ERROR, "internal: unexpected stride pattern");
      if ((op[k].kind == K_FULL || op[k].kind == K_ROW) && !aligned16(ins[k].p)) vec_ok = false;
      if (op[k].kind == K_ROW || op[k].kind == K_COL) need_rc = true;
    }
    // rows == 1: vectorise the first floor(cols/4)*4 elements, scalar tail.
    // rows > 1: vectorise everything when cols % 4 == 0.
    uint64_t vec_elems = 0;
    if (vec_ok) {
      if (rows == 1) vec_elems = (cols / 4) * 4;
      else if (cols % 4 == 0) vec_elems = rows * cols;
    }
    const uint64_t total4 = vec_elems / 4;
    if (total4 > 0) {
      const uint64_t cols4 = rows == 1 ? total4 : cols / 4;
      if (total4 < (uint64_t(1) << 31)) {
        unsigned grid = full_grid(total4, EW_BLOCK);
        FastDiv fd((uint32_t)cols4);
        if (launch_kind(dim3(grid), g, f, op, out, (uint32_t)total4, fd))
          ;  // specialised kernel
        else if (need_rc)
          ew_launch(g, k_ew2d_vec<F, true>, dim3(grid), f, op[0], op[1], op[2], out, (uint32_t)total4, fd);
        else
          ew_launch(g, k_ew2d_vec<F, false>, dim3(grid), f, op[0], op[1], op[2], out, (uint32_t)total4, fd);
      } else {
        unsigned grid = stream_grid(g, total4, EW_BLOCK, 8);
        ew_launch(g, k_ew2d_vec64<F>, dim3(grid), f, op[0], op[1], op[2], out, total4, cols4);
      }
      check_launch("elementwise (vector)");
    }
    if (vec_elems < pl.total) {
      unsigned grid = stream_grid(g, pl.total - vec_elems, EW_BLOCK, 8);
      ew_launch(g, k_ew2d_scalar<F>, dim3(grid), f, op[0], op[1], op[2], out, vec_elems, pl.total, cols);
      check_launch("elementwise (scalar)");
    }
    return;
  }
  VKH_REQUIRE(nd <= MAXD, "too many dimensions after collapsing");
  NdDesc d{};
  d.nd = nd;
  for (int k = 0; k < nd; ++k) d.dims[k] = pl.dims[k];
  for (size_t k = 0; k < 3; ++k) {
    d.p[k] = k < nin ? ins[k].p : nullptr;
    d.s[k] = k < nin ? ins[k].s : 0.f;
    for (int j = 0; j < nd; ++j) d.st[k][j] = (k < nin && ins[k].p) ? pl.st[k][j] : 0;
  }
  unsigned grid = stream_grid(g, pl.total, EW_BLOCK, 8);
  ew_launch(g, k_ewnd<F>, dim3(grid), f, d, out, pl.total);
  check_launch("elementwise (n-d)");
}

static std::vector<uint64_t> to_vec(const uint64_t* s, int ndim) {
  VKH_REQUIRE(ndim >= 0 && ndim <= 64, "bad ndim");
  VKH_REQUIRE(ndim == 0 || s != nullptr, "null shape");
  return std::vector<uint64_t>(s, s + ndim);
}

template <int OP>
struct BinRun {
  static void run(vkh_gpu* g, const std::vector<InSpec>& ins, float* out,
                  const std::vector<uint64_t>& oshape) {
    run_ew(g, BinF<OP>{}, ins, out, oshape);
  }
};

// ---- unary math -------------------------------------------------------------------
template <int FN>
__device__ __forceinline__ float unop(float x) {
  if constexpr (FN == VKH_ABS) return fabsf(x);
  else if constexpr (FN == VKH_SIGN) return x > 0.f ? 1.f : (x < 0.f ? -1.f : (x == 0.f ? 0.f : x));
  else if constexpr (FN == VKH_SIN) return sin_f32(x);
  else if constexpr (FN == VKH_COS) return cos_f32(x);
  else if constexpr (FN == VKH_TAN) return tan_f32(x);
  else if constexpr (FN == VKH_ASIN) return asin_f32(x);
  else if constexpr (FN == VKH_ACOS) return acos_f32(x);
  else if constexpr (FN == VKH_ATAN) return atan_f32(x);
  else if constexpr (FN == VKH_SINH) return sinh_f32(x);
  else if constexpr (FN == VKH_COSH) return cosh_f32(x);
  else if constexpr (FN == VKH_TANH) return tanh_f32(x);
  else if constexpr (FN == VKH_ASINH) return asinh_f32(x);
  else if constexpr (FN == VKH_ACOSH) return acosh_f32(x);
  else if constexpr (FN == VKH_ATANH) return atanh_f32(x);
  else if constexpr (FN == VKH_EXP) return exp_f32(x);
  else if constexpr (FN == VKH_LOG) return log_f32(x);
  else if constexpr (FN == VKH_EXP2) return exp2_f32(x);
  else if constexpr (FN == VKH_LOG2) return log2_f32(x);
  else if constexpr (FN == VKH_SQRT) return sqrtf(x);
  else return invsqrt_f32(x);
}

template <int FN>
__global__ void __launch_bounds__(EW_BLOCK)
k_unary_vec(const float4* a, float4* b, uint64_t n4) {
  const uint64_t stride = (uint64_t)gridDim.x * EW_BLOCK;
  for (uint64_t t = (uint64_t)blockIdx.x * EW_BLOCK + threadIdx.x; t < n4; t += stride) {
    const float4 v = ld_stream(a + t);
    st_stream(b + t, make_float4(unop<FN>(v.x), unop<FN>(v.y), unop<FN>(v.z), unop<FN>(v.w)));
  }
}

template <int FN>
__global__ void __launch_bounds__(EW_BLOCK)
k_unary(const float* a, float* b, uint64_t begin, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * EW_BLOCK;
  for (uint64_t t = begin + (uint64_t)blockIdx.x * EW_BLOCK + threadIdx.x; t < n; t += stride)
    b[t] = unop<FN>(a[t]);
}

template <int FN>
static void run_unary(vkh_gpu* g, const float* a, float* b, uint64_t n) {
  uint64_t vec = 0;
  if (aligned16(a) && aligned16(b)) {
    uint64_t n4 = n / 4;
    vec = n4 * 4;
    if (n4) {
      unsigned grid = full_grid(n4, EW_BLOCK);
      ew_launch(g, k_unary_vec<FN>, dim3(grid), (const float4*)a, (float4*)b, n4);
      check_launch("unary (vector)");
    }
  }
  if (vec < n) {
    unsigned grid = stream_grid(g, n - vec, EW_BLOCK, 8);
    ew_launch(g, k_unary<FN>, dim3(grid), a, b, vec, n);
    check_launch("unary");
  }
}

static void dispatch_unary(vkh_gpu* g, int fn, const float* a, float* b, uint64_t n) {
  switch (fn) {
#define U(X) case X: run_unary<X>(g, a, b, n); break;
    U(VKH_ABS) U(VKH_SIGN) U(VKH_SIN) U(VKH_COS) U(VKH_TAN) U(VKH_ASIN) U(VKH_ACOS)
    U(VKH_ATAN) U(VKH_SINH) U(VKH_COSH) U(VKH_TANH) U(VKH_ASINH) U(VKH_ACOSH)
    U(VKH_ATANH) U(VKH_EXP) U(VKH_LOG) U(VKH_EXP2) U(VKH_LOG2) U(VKH_SQRT) U(VKH_INVSQRT)
#undef U
    default: fail(VKH_INVALID, "unknown unary function " + std::to_string(fn));
  }
}

}  // namespace vkh

using namespace vkh;

extern "C" {

int vkh_binary(vkh_gpu* g, int op, const float* a, const float* b, float* c, uint64_t n) {
  return guarded([&] {
    use(g);
    if (n == 0) return;
    VKH_REQUIRE(a && b && c, "binary: null pointer");
    std::vector<uint64_t> s{n};
    std::vector<InSpec> ins{{a, 0.f, s}, {b, 0.f, s}};
    dispatch_binop<BinRun>(op, g, ins, c, s);
  });
}

int vkh_binary_scalar(vkh_gpu* g, int op, const float* a, float sv, float* c, uint64_t n) {
  return guarded([&] {
    use(g);
    if (n == 0) return;
    VKH_REQUIRE(a && c, "binary_scalar: null pointer");
    std::vector<uint64_t> s{n};
    std::vector<InSpec> ins{{a, 0.f, s}, {nullptr, sv, {}}};
    if (op == VKH_POW && (sv == 1.0f || sv == 2.0f || sv == -1.0f || sv == 0.0f || sv == 0.5f)) {
      if (sv == 1.0f) run_ew(g, PowScalarF<1>{}, ins, c, s);
      else if (sv == 2.0f) run_ew(g, PowScalarF<2>{}, ins, c, s);
      else if (sv == -1.0f) run_ew(g, PowScalarF<-1>{}, ins, c, s);
      else if (sv == 0.0f) run_ew(g, PowScalarF<0>{}, ins, c, s);
      else run_ew(g, PowScalarF<5>{}, ins, c, s);
      return;
    }
    dispatch_binop<BinRun>(op, g, ins, c, s);
  });
}

int vkh_binary_bcast(vkh_gpu* g, int op, const float* a, const float* b, float* c, int ndim,
                     const uint64_t* sa, const uint64_t* sb, const uint64_t* sc) {
  return guarded([&] {
    use(g);
    VKH_REQUIRE(a && b && c, "binary_bcast: null pointer");
    auto va = to_vec(sa, ndim), vb = to_vec(sb, ndim), vc = to_vec(sc, ndim);
    if (c == a) {
      for (int d = 0; d < ndim; ++d)
        VKH_REQUIRE(va[d] == vc[d], "in-place broadcast: output shape must equal the first operand's");
    }
    std::vector<InSpec> ins{{a, 0.f, va}, {b, 0.f, vb}};
    dispatch_binop<BinRun>(op, g, ins, c, vc);
  });
}

int vkh_unary(vkh_gpu* g, int fn, const float* a, float* b, uint64_t n) {
  return guarded([&] {
    use(g);
    if (n == 0) return;
    VKH_REQUIRE(a && b, "unary: null pointer");
    dispatch_unary(g, fn, a, b, n);
  });
}

int vkh_clamp(vkh_gpu* g, const float* x, const float* lo, const float* hi, float lo_s,
              float hi_s, float* out, int ndim, const uint64_t* sx, const uint64_t* slo,
              const uint64_t* shi, const uint64_t* so) {
  return guarded([&] {
    use(g);
    VKH_REQUIRE(x && out, "clamp: null pointer");
    auto vo = to_vec(so, ndim);
    std::vector<InSpec> ins{{x, 0.f, to_vec(sx, ndim)},
                            {lo, lo_s, lo ? to_vec(slo, ndim) : std::vector<uint64_t>{}},
                            {hi, hi_s, hi ? to_vec(shi, ndim) : std::vector<uint64_t>{}}};
    run_ew(g, ClampF{}, ins, out, vo);
  });
}

int vkh_broadcast_to(vkh_gpu* g, const float* a, float* out, int ndim, const uint64_t* sa,
                     const uint64_t* so) {
  return guarded([&] {
    use(g);
    VKH_REQUIRE(a && out, "broadcast_to: null pointer");
    std::vector<InSpec> ins{{a, 0.f, to_vec(sa, ndim)}};
    run_ew(g, CopyF{}, ins, out, to_vec(so, ndim));
  });
}

}  // extern "C"
