// Fused kernels of the nn path: activations and their backward, row softmax,
// cross entropy (F19), MSE / Huber losses, loss gradients, optimizers and the
// Dense bias gradient.
//
// Each kernel performs, per element, the same FP32 operations in the same
// order as the Array-op composition in the reference's nn package (cited per
// kernel), so the fused result equals the composed one bit for bit; row sums
// over short rows are sequential like vulkpy/shader/sum_axis.comp:28-30.
#include "common.h"

#include <initializer_list>

namespace vkh {

__device__ __forceinline__ float sign_np(float x) {
  return x > 0.f ? 1.f : (x < 0.f ? -1.f : (x == 0.f ? 0.f : x));
}

// nn/layers.py:185 (x.max(0.0)) and 240-244 (0-x, exp, +1, 1/y)
__device__ __forceinline__ float act_fwd(int act, float x) {
  if (act == VKH_ACT_RELU) return np_max(x, 0.0f);
  if (act == VKH_ACT_SIGMOID) {
    float y = 0.0f - x;
    y = dexp(y);
    y = y + 1.0f;
    return 1.0f / y;
  }
  return x;
}

// nn/layers.py:207-210 and 264-267
__device__ __forceinline__ float act_bwd(int act, float y, float dy) {
  if (act == VKH_ACT_RELU) {
    float d = np_max(sign_np(y), 0.0f);
    return d * dy;
  }
  if (act == VKH_ACT_SIGMOID) {
    float d = 1.0f - y;
    d = d * y;
    return d * dy;
  }
  return dy;
}

__global__ void __launch_bounds__(256) k_act(int act, const float* x, float* y, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) y[i] = act_fwd(act, x[i]);
}

__global__ void __launch_bounds__(256) k_act_bwd(int act, const float* y, const float* dy, float* dx, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) dx[i] = act_bwd(act, y[i], dy[i]);
}

// ---- row softmax and cross entropy --------------------------------------------------
// exp(x - max) with FP32 denormal results flushed to zero: the reference runs
// on Vulkan drivers that flush denormals, and its own test expects
// softmax([100, 0]) == [1, 0] exactly (test/test_nn.py:124-131), where the
// IEEE result of exp(-100) is the denormal 3.8e-44.
__device__ __forceinline__ float softmax_exp(float v) {
  const float e = dexp(v);
  return e < 1.17549435e-38f ? 0.0f : e;
}

// nn/layers.py:297-300: X = x - max(x); X = exp(X); X /= sum(X)
__global__ void __launch_bounds__(256)
k_softmax_rows_thread(const float* x, float* y, uint64_t rows, uint64_t cols) {
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t r = (uint64_t)blockIdx.x * 256 + threadIdx.x; r < rows; r += stride) {
    const float* xr = x + r * cols;
    float* yr = y + r * cols;
    float m = xr[0];
    for (uint64_t j = 0; j < cols; ++j) m = np_max(m, xr[j]);
    float s = 0.0f;
    for (uint64_t j = 0; j < cols; ++j) {
      float e = softmax_exp(xr[j] - m);
      yr[j] = e;
      s += e;
    }
    for (uint64_t j = 0; j < cols; ++j) yr[j] = yr[j] / s;
  }
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = np_max(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ float block_max(float v, float* lds) {
  v = wave_max(v);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
  __syncthreads();
  float r = lds[0];
  for (int i = 1; i < 4; ++i) r = np_max(r, lds[i]);
  __syncthreads();
  return r;
}
__device__ __forceinline__ float block_sum(float v, float* lds) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
  __syncthreads();
  float r = lds[0] + lds[1] + lds[2] + lds[3];
  __syncthreads();
  return r;
}

__global__ void __launch_bounds__(256)
k_softmax_rows_block(const float* x, float* y, uint64_t rows, uint64_t cols) {
  __shared__ float lds[4];
  for (uint64_t r = blockIdx.x; r < rows; r += gridDim.x) {
    const float* xr = x + r * cols;
    float* yr = y + r * cols;
    float m = -__builtin_huge_valf();
    for (uint64_t j = threadIdx.x; j < cols; j += 256) m = np_max(m, xr[j]);
    m = block_max(m, lds);
    float s = 0.0f;
    for (uint64_t j = threadIdx.x; j < cols; j += 256) {
      float e = softmax_exp(xr[j] - m);
      yr[j] = e;
      s += e;
    }
    s = block_sum(s, lds);
    for (uint64_t j = threadIdx.x; j < cols; j += 256) yr[j] = yr[j] / s;
  }
}

// Softmax Jacobian-vector product: dx = y * (dy - sum_k dy[k] y[k]) per row
// (the exact gradient, used by the autograd wrapper; the reference layer's
// own backward is the diagonal form dy * y * (1 - y), k_act_bwd's job).
__global__ void __launch_bounds__(256)
k_softmax_bwd_thread(const float* y, const float* dy, float* dx, uint64_t rows, uint64_t cols) {
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t r = (uint64_t)blockIdx.x * 256 + threadIdx.x; r < rows; r += stride) {
    const float* yr = y + r * cols;
    const float* gr = dy + r * cols;
    float s = 0.0f;
    for (uint64_t j = 0; j < cols; ++j) s += gr[j] * yr[j];
    for (uint64_t j = 0; j < cols; ++j) dx[r * cols + j] = yr[j] * (gr[j] - s);
  }
}

__global__ void __launch_bounds__(256)
k_softmax_bwd_block(const float* y, const float* dy, float* dx, uint64_t rows, uint64_t cols) {
  __shared__ float lds[4];
  for (uint64_t r = blockIdx.x; r < rows; r += gridDim.x) {
    const float* yr = y + r * cols;
    const float* gr = dy + r * cols;
    float s = 0.0f;
    for (uint64_t j = threadIdx.x; j < cols; j += 256) s += gr[j] * yr[j];
    s = block_sum(s, lds);
    for (uint64_t j = threadIdx.x; j < cols; j += 256) dx[r * cols + j] = yr[j] * (gr[j] - s);
  }
}

// nn_cross_entropy.comp:26 + sum(axis=1): L[r] = sum_j (-y) * log(x + 1e-8)
__device__ __forceinline__ float xent_term(float x, float y) { return (-y) * dlog(x + 1e-8f); }

__global__ void __launch_bounds__(256)
k_xent_rows(const float* x, const float* y, float* L, uint64_t rows, uint64_t cols) {
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t r = (uint64_t)blockIdx.x * 256 + threadIdx.x; r < rows; r += stride) {
    float acc = 0.0f;
    for (uint64_t j = 0; j < cols; ++j) acc += xent_term(x[r * cols + j], y[r * cols + j]);
    L[r] = acc;
  }
}

// softmax + cross entropy per row (nn/losses.py:227), sequential rows
__global__ void __launch_bounds__(256)
k_sm_xent_rows(const float* x, const float* y, float* sm, float* L, uint64_t rows, uint64_t cols) {
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t r = (uint64_t)blockIdx.x * 256 + threadIdx.x; r < rows; r += stride) {
    const float* xr = x + r * cols;
    const float* yr = y + r * cols;
    float* sr = sm + r * cols;
    float m = xr[0];
    for (uint64_t j = 0; j < cols; ++j) m = np_max(m, xr[j]);
    float s = 0.0f;
    for (uint64_t j = 0; j < cols; ++j) {
      float e = softmax_exp(xr[j] - m);
      sr[j] = e;
      s += e;
    }
    float acc = 0.0f;
    for (uint64_t j = 0; j < cols; ++j) {
      float p = sr[j] / s;
      sr[j] = p;
      acc += xent_term(p, yr[j]);
    }
    L[r] = acc;
  }
}

// nn/losses.py:294-296 (MSE) and 365-369 (Huber), per row, sequential sum
template <int KIND>
__global__ void __launch_bounds__(256)
k_loss_rows(const float* x, const float* y, float* L, uint64_t rows, uint64_t cols) {
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t r = (uint64_t)blockIdx.x * 256 + threadIdx.x; r < rows; r += stride) {
    float acc = 0.0f;
    for (uint64_t j = 0; j < cols; ++j) {
      float d = y[r * cols + j] - x[r * cols + j];
      float t;
      if (KIND == 0) {
        t = d * d;  // (y-x) ** 2.0, exact square rounded once
      } else {
        float a = fabsf(d);
        t = np_min(a, a * a);
        t = t * 0.5f;
      }
      acc += t;
    }
    L[r] = acc;
  }
}

// nn/losses.py:249, 318-320, 391-393 then ReduceLoss.grad's `dx *= scale`
__global__ void __launch_bounds__(256)
k_loss_grad(int kind, const float* x, const float* y, float* dx, uint64_t n, float scale) {
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
    float d = x[i] - y[i];
    if (kind == 1) d = d * 2.0f;
    else if (kind == 2) d = np_min(np_max(d, -1.0f), 1.0f);
    else if (kind == 3) d = (fabsf(d) <= 1.0f || d != d) ? d : (d > 0.0f ? 0.5f : -0.5f);
    if (scale != 1.0f) d = d * scale;
    dx[i] = d;
  }
}

// k_loss_rows<KIND> and k_loss_grad (kind 1 for MSE, 2 for Huber) in one
// pass over x and y: the same bits, one launch fewer
template <int KIND>
__global__ void __launch_bounds__(256)
k_loss_rows_grad(const float* x, const float* y, float* L, float* dx, uint64_t rows, uint64_t cols,
                 float scale) {
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t r = (uint64_t)blockIdx.x * 256 + threadIdx.x; r < rows; r += stride) {
    float acc = 0.0f;
    for (uint64_t j = 0; j < cols; ++j) {
      const uint64_t i = r * cols + j;
      const float xv = x[i], yv = y[i];
      const float d = yv - xv;
      float t;
      if (KIND == 0) {
        t = d * d;
      } else {
        const float a = fabsf(d);
        t = np_min(a, a * a);
        t = t * 0.5f;
      }
      acc += t;
      float gd = xv - yv;
      gd = KIND == 0 ? gd * 2.0f : np_min(np_max(gd, -1.0f), 1.0f);
      if (scale != 1.0f) gd = gd * scale;
      dx[i] = gd;
    }
    L[r] = acc;
  }
}

// nn_cross_entropy_backward.comp:26 then `dx *= scale`
__global__ void __launch_bounds__(256)
k_xent_bwd(const float* x, const float* y, float* dx, uint64_t n, float scale) {
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
    float d = (-y[i]) / (x[i] + 1e-8f);
    if (scale != 1.0f) d = d * scale;
    dx[i] = d;
  }
}

// ---- regularizer gradients (nn/regularizers.py:23-192) ------------------------------
// kind 1 Lasso.grad  coeff * sign(w)
// kind 2 Ridge.grad  (2 coeff) * w, the scalar 2 coeff formed in FP64 and rounded
// kind 3 Elastic.grad  L1.grad + L2.grad, each rounded before the sum
// Parameter.regular_grad then adds the term onto the gradient: g + r.
struct RegTerms {
  int kind;  // 0: no regularizer
  float l1, s2;
};

__device__ __forceinline__ float reg_term(const RegTerms& r, float w) {
  const float t1 = sign_np(w) * r.l1;
  const float t2 = w * r.s2;
  return r.kind == 1 ? t1 : (r.kind == 2 ? t2 : t1 + t2);
}

__device__ __forceinline__ float with_reg(const RegTerms& r, float g, float w) {
  return r.kind ? g + reg_term(r, w) : g;
}

// Element loops of the parameter kernels: 16-byte accesses over the first
// 4 * n4 elements (n4 = 0 when a pointer is not 16-byte aligned), then the
// scalar tail.
template <typename F4, typename F1>
__device__ __forceinline__ void vec_loop(uint64_t n4, uint64_t n, F4&& body4, F1&& body1) {
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  const uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  for (uint64_t q = t; q < n4; q += stride) body4(q);
  for (uint64_t i = 4 * n4 + t; i < n; i += stride) body1(i);
}

__global__ void __launch_bounds__(256)
k_reg_grad(RegTerms r, const float* __restrict__ w, float* __restrict__ g, uint64_t n, uint64_t n4,
           int accumulate) {
  vec_loop(n4, n, [&](uint64_t q) {
    const float4 wv = reinterpret_cast<const float4*>(w)[q];
    float4 gv = accumulate ? reinterpret_cast<float4*>(g)[q] : make_float4(0.f, 0.f, 0.f, 0.f);
    gv.x = accumulate ? gv.x + reg_term(r, wv.x) : reg_term(r, wv.x);
    gv.y = accumulate ? gv.y + reg_term(r, wv.y) : reg_term(r, wv.y);
    gv.z = accumulate ? gv.z + reg_term(r, wv.z) : reg_term(r, wv.z);
    gv.w = accumulate ? gv.w + reg_term(r, wv.w) : reg_term(r, wv.w);
    reinterpret_cast<float4*>(g)[q] = gv;
  }, [&](uint64_t i) {
    const float t = reg_term(r, w[i]);
    g[i] = accumulate ? g[i] + t : t;
  });
}

// ---- optimizers ----------------------------------------------------------------------
struct AdamScalars {
  float beta1, one_m_b1, beta2, one_m_b2, eps, neg_lr, c1, c2;
};

// nn/optimizers.py:235-253 and nn/parameters.py:95, in that op order
__device__ __forceinline__ void adam_elem(float& w, float g, float& m, float& v, const AdamScalars& s) {
  m = m * s.beta1;
  m = m + g * s.one_m_b1;
  v = v * s.beta2;
  v = v + (g * g) * s.one_m_b2;
  float mhat = m / s.c1;
  float vhat = v / s.c2;
  vhat = sqrtf(vhat);
  vhat = vhat + s.eps;
  mhat = mhat * s.neg_lr;
  mhat = mhat / vhat;
  w = w + mhat;
}


// One Adam step of parameter element(s): the regularizer term (if any) is
// added to the gradient first and the sum written back, as
// Parameter.regular_grad leaves it, then the update reads it.
__device__ __forceinline__ void adam4(float4& w, float4& g, float4& m, float4& v, const AdamScalars& s,
                                      const RegTerms& r) {
  g.x = with_reg(r, g.x, w.x); g.y = with_reg(r, g.y, w.y);
  g.z = with_reg(r, g.z, w.z); g.w = with_reg(r, g.w, w.w);
  adam_elem(w.x, g.x, m.x, v.x, s); adam_elem(w.y, g.y, m.y, v.y, s);
  adam_elem(w.z, g.z, m.z, v.z, s); adam_elem(w.w, g.w, m.w, v.w, s);
}

__global__ void __launch_bounds__(256)
k_adam(float* __restrict__ w, float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
       uint64_t n, uint64_t n4, AdamScalars s, const double* bpow, RegTerms r, uint16_t* wbf) {
  if (bpow) {  // beta powers kept on the device (captured steps)
    s.c1 = (float)(1.0 - bpow[0]);
    s.c2 = (float)(1.0 - bpow[1]);
  }
  vec_loop(n4, n, [&](uint64_t q) {
    float4 wv = reinterpret_cast<float4*>(w)[q], gv = reinterpret_cast<const float4*>(g)[q];
    float4 mv = reinterpret_cast<float4*>(m)[q], vv = reinterpret_cast<float4*>(v)[q];
    adam4(wv, gv, mv, vv, s, r);
    reinterpret_cast<float4*>(w)[q] = wv;
    reinterpret_cast<float4*>(m)[q] = mv;
    reinterpret_cast<float4*>(v)[q] = vv;
    if (r.kind) reinterpret_cast<float4*>(g)[q] = gv;
    if (wbf) {
      ushort4 b;
      b.x = f2bf(wv.x); b.y = f2bf(wv.y); b.z = f2bf(wv.z); b.w = f2bf(wv.w);
      reinterpret_cast<ushort4*>(wbf)[q] = b;
    }
  }, [&](uint64_t i) {
    float wi = w[i], gi = with_reg(r, g[i], w[i]), mi = m[i], vi = v[i];
    adam_elem(wi, gi, mi, vi, s);
    w[i] = wi; m[i] = mi; v[i] = vi;
    if (r.kind) g[i] = gi;
    if (wbf) wbf[i] = f2bf(wi);
  });
}

// Beta powers live in device memory: bpow[0] = beta1^t, bpow[1] = beta2^t
// (FP64 like the Python floats of AdamState).  k_adam_pow advances them once.
__global__ void k_adam_pow(double* bpow, double beta1, double beta2) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    bpow[0] = bpow[0] * beta1;
    bpow[1] = bpow[1] * beta2;
  }
}

// Several parameters' Adam updates in one launch: parameter p takes blocks
// [first[p], first[p + 1]) and walks its elements as k_adam does with that
// many blocks (the element arithmetic is k_adam's, so the bits are too).
constexpr int ADAM_MT = VKH_ADAM_MULTI_MAX;
struct AdamTensor {
  float *w, *g, *m, *v;
  uint64_t n, n4;
  AdamScalars s;
  const double* bpow;
  RegTerms r;
};
struct AdamMulti {
  AdamTensor t[ADAM_MT];
  uint32_t first[ADAM_MT + 1];
  int count;
};

__global__ void __launch_bounds__(256) k_adam_multi(AdamMulti a) {
  // the block's parameter, by constant indices only (a by-value argument
  // indexed at run time would be copied to scratch memory)
  AdamTensor t = a.t[0];
  uint32_t b0 = a.first[0], b1 = a.first[1];
#pragma unroll
  for (int q = 1; q < ADAM_MT; ++q)
    if (q < a.count && blockIdx.x >= a.first[q]) {
      t = a.t[q];
      b0 = a.first[q];
      b1 = a.first[q + 1];
    }
  AdamScalars s = t.s;
  if (t.bpow) {
    s.c1 = (float)(1.0 - t.bpow[0]);
    s.c2 = (float)(1.0 - t.bpow[1]);
  }
  float* __restrict__ w = t.w;
  float* __restrict__ g = t.g;
  float* __restrict__ m = t.m;
  float* __restrict__ v = t.v;
  const RegTerms r = t.r;
  const uint64_t stride = (uint64_t)(b1 - b0) * 256;
  const uint64_t i0 = (uint64_t)(blockIdx.x - b0) * 256 + threadIdx.x;
  for (uint64_t q = i0; q < t.n4; q += stride) {
    float4 wv = reinterpret_cast<float4*>(w)[q], gv = reinterpret_cast<const float4*>(g)[q];
    float4 mv = reinterpret_cast<float4*>(m)[q], vv = reinterpret_cast<float4*>(v)[q];
    adam4(wv, gv, mv, vv, s, r);
    reinterpret_cast<float4*>(w)[q] = wv;
    reinterpret_cast<float4*>(m)[q] = mv;
    reinterpret_cast<float4*>(v)[q] = vv;
    if (r.kind) reinterpret_cast<float4*>(g)[q] = gv;
  }
  for (uint64_t i = 4 * t.n4 + i0; i < t.n; i += stride) {
    float wi = w[i], gi = with_reg(r, g[i], w[i]), mi = m[i], vi = v[i];
    adam_elem(wi, gi, mi, vi, s);
    w[i] = wi; m[i] = mi; v[i] = vi;
    if (r.kind) g[i] = gi;
  }
}

// nn/optimizers.py:53: w += (-lr) * g (g with the regularizer term added first)
__global__ void __launch_bounds__(256)
k_sgd(float* __restrict__ w, float* __restrict__ g, uint64_t n, uint64_t n4, float neg_lr, RegTerms r) {
  vec_loop(n4, n, [&](uint64_t q) {
    float4 wv = reinterpret_cast<float4*>(w)[q], gv = reinterpret_cast<const float4*>(g)[q];
    gv.x = with_reg(r, gv.x, wv.x); gv.y = with_reg(r, gv.y, wv.y);
    gv.z = with_reg(r, gv.z, wv.z); gv.w = with_reg(r, gv.w, wv.w);
    wv.x = wv.x + gv.x * neg_lr; wv.y = wv.y + gv.y * neg_lr;
    wv.z = wv.z + gv.z * neg_lr; wv.w = wv.w + gv.w * neg_lr;
    reinterpret_cast<float4*>(w)[q] = wv;
    if (r.kind) reinterpret_cast<float4*>(g)[q] = gv;
  }, [&](uint64_t i) {
    const float gi = with_reg(r, g[i], w[i]);
    w[i] = w[i] + gi * neg_lr;
    if (r.kind) g[i] = gi;
  });
}

// nn/optimizers.py:134-140
__device__ __forceinline__ void adagrad_elem(float& w, float g, float& h, float neg_lr, float eps) {
  h = h + g * g;
  float s = sqrtf(h);
  s = s + eps;
  float q = g / s;
  q = q * neg_lr;
  w = w + q;
}

__global__ void __launch_bounds__(256)
k_adagrad(float* __restrict__ w, float* __restrict__ g, float* __restrict__ h, uint64_t n, uint64_t n4,
          float neg_lr, float eps, RegTerms r) {
  vec_loop(n4, n, [&](uint64_t q) {
    float4 wv = reinterpret_cast<float4*>(w)[q], gv = reinterpret_cast<const float4*>(g)[q];
    float4 hv = reinterpret_cast<float4*>(h)[q];
    gv.x = with_reg(r, gv.x, wv.x); gv.y = with_reg(r, gv.y, wv.y);
    gv.z = with_reg(r, gv.z, wv.z); gv.w = with_reg(r, gv.w, wv.w);
    adagrad_elem(wv.x, gv.x, hv.x, neg_lr, eps); adagrad_elem(wv.y, gv.y, hv.y, neg_lr, eps);
    adagrad_elem(wv.z, gv.z, hv.z, neg_lr, eps); adagrad_elem(wv.w, gv.w, hv.w, neg_lr, eps);
    reinterpret_cast<float4*>(w)[q] = wv;
    reinterpret_cast<float4*>(h)[q] = hv;
    if (r.kind) reinterpret_cast<float4*>(g)[q] = gv;
  }, [&](uint64_t i) {
    float wi = w[i], gi = with_reg(r, g[i], w[i]), hi = h[i];
    adagrad_elem(wi, gi, hi, neg_lr, eps);
    w[i] = wi; h[i] = hi;
    if (r.kind) g[i] = gi;
  });
}

__global__ void __launch_bounds__(256)
k_axpby(float* db, const float* t, uint64_t n, float beta) {
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride)
    db[i] = beta == 0.0f ? t[i] : db[i] * beta + t[i];
}

}  // namespace vkh

using namespace vkh;

static unsigned grid_for(vkh_gpu* g, uint64_t n) { return stream_grid(g, n, 256, 8); }

extern "C" {

int vkh_activation(vkh_gpu* g, int act, const float* x, float* y, uint64_t n) {
  return guarded([&] {
    use(g);
    if (n == 0) return;
    VKH_REQUIRE(x && y, "activation: null pointer");
    hipLaunchKernelGGL(k_act, dim3(grid_for(g, n)), dim3(256), 0, g->stream, act, x, y, n);
    check_launch("activation");
  });
}

int vkh_activation_backward(vkh_gpu* g, int act, const float* y, const float* dy, float* dx, uint64_t n) {
  return guarded([&] {
    use(g);
    if (n == 0) return;
    VKH_REQUIRE(y && dy && dx, "activation_backward: null pointer");
    hipLaunchKernelGGL(k_act_bwd, dim3(grid_for(g, n)), dim3(256), 0, g->stream, act, y, dy, dx, n);
    check_launch("activation_backward");
  });
}

int vkh_softmax(vkh_gpu* g, const float* x, float* y, uint64_t rows, uint64_t cols) {
  return guarded([&] {
    use(g);
    if (rows == 0 || cols == 0) return;
    VKH_REQUIRE(x && y, "softmax: null pointer");
    if (cols <= 64) {
      unsigned grid = (unsigned)std::min<uint64_t>(ceil_div(rows, 256), (uint64_t)g->num_cus * 8);
      hipLaunchKernelGGL(k_softmax_rows_thread, dim3(grid), dim3(256), 0, g->stream, x, y, rows, cols);
    } else {
      unsigned grid = (unsigned)std::min<uint64_t>(rows, (uint64_t)g->num_cus * 16);
      hipLaunchKernelGGL(k_softmax_rows_block, dim3(grid), dim3(256), 0, g->stream, x, y, rows, cols);
    }
    check_launch("softmax");
  });
}

int vkh_softmax_backward(vkh_gpu* g, const float* y, const float* dy, float* dx, uint64_t rows,
                         uint64_t cols) {
  return guarded([&] {
    use(g);
    if (rows == 0 || cols == 0) return;
    VKH_REQUIRE(y && dy && dx, "softmax_backward: null pointer");
    if (cols <= 64) {
      unsigned grid = (unsigned)std::min<uint64_t>(ceil_div(rows, 256), (uint64_t)g->num_cus * 8);
      hipLaunchKernelGGL(k_softmax_bwd_thread, dim3(grid), dim3(256), 0, g->stream, y, dy, dx, rows, cols);
    } else {
      unsigned grid = (unsigned)std::min<uint64_t>(rows, (uint64_t)g->num_cus * 16);
      hipLaunchKernelGGL(k_softmax_bwd_block, dim3(grid), dim3(256), 0, g->stream, y, dy, dx, rows, cols);
    }
    check_launch("softmax_backward");
  });
}

int vkh_cross_entropy(vkh_gpu* g, const float* x, const float* y, float* L, uint64_t rows, uint64_t cols) {
  return guarded([&] {
    use(g);
    if (rows == 0) return;
    VKH_REQUIRE(x && y && L, "cross_entropy: null pointer");
    unsigned grid = (unsigned)std::min<uint64_t>(ceil_div(rows, 256), (uint64_t)g->num_cus * 8);
    hipLaunchKernelGGL(k_xent_rows, dim3(grid), dim3(256), 0, g->stream, x, y, L, rows, cols);
    check_launch("cross_entropy");
  });
}

int vkh_cross_entropy_backward(vkh_gpu* g, const float* x, const float* y, float* dx, uint64_t n, float scale) {
  return guarded([&] {
    use(g);
    if (n == 0) return;
    VKH_REQUIRE(x && y && dx, "cross_entropy_backward: null pointer");
    hipLaunchKernelGGL(k_xent_bwd, dim3(grid_for(g, n)), dim3(256), 0, g->stream, x, y, dx, n, scale);
    check_launch("cross_entropy_backward");
  });
}

int vkh_softmax_cross_entropy(vkh_gpu* g, const float* x, const float* y, float* sm, float* L,
                              uint64_t rows, uint64_t cols) {
  return guarded([&] {
    use(g);
    if (rows == 0 || cols == 0) return;
    VKH_REQUIRE(x && y && sm && L, "softmax_cross_entropy: null pointer");
    if (cols <= 64) {
      unsigned grid = (unsigned)std::min<uint64_t>(ceil_div(rows, 256), (uint64_t)g->num_cus * 8);
      hipLaunchKernelGGL(k_sm_xent_rows, dim3(grid), dim3(256), 0, g->stream, x, y, sm, L, rows, cols);
      check_launch("softmax_cross_entropy");
      return;
    }
    int r = vkh_softmax(g, x, sm, rows, cols);
    if (r != VKH_OK) fail(r, vkh_last_error());
    r = vkh_cross_entropy(g, sm, y, L, rows, cols);
    if (r != VKH_OK) fail(r, vkh_last_error());
  });
}

int vkh_mse_rows(vkh_gpu* g, const float* x, const float* y, float* L, uint64_t rows, uint64_t cols) {
  return guarded([&] {
    use(g);
    if (rows == 0) return;
    VKH_REQUIRE(x && y && L, "mse_rows: null pointer");
    unsigned grid = (unsigned)std::min<uint64_t>(ceil_div(rows, 256), (uint64_t)g->num_cus * 8);
    hipLaunchKernelGGL(k_loss_rows<0>, dim3(grid), dim3(256), 0, g->stream, x, y, L, rows, cols);
    check_launch("mse_rows");
  });
}

int vkh_huber_rows(vkh_gpu* g, const float* x, const float* y, float* L, uint64_t rows, uint64_t cols) {
  return guarded([&] {
    use(g);
    if (rows == 0) return;
    VKH_REQUIRE(x && y && L, "huber_rows: null pointer");
    unsigned grid = (unsigned)std::min<uint64_t>(ceil_div(rows, 256), (uint64_t)g->num_cus * 8);
    hipLaunchKernelGGL(k_loss_rows<1>, dim3(grid), dim3(256), 0, g->stream, x, y, L, rows, cols);
    check_launch("huber_rows");
  });
}

int vkh_loss_rows_grad(vkh_gpu* g, int kind, const float* x, const float* y, float* L, float* dx,
                       uint64_t rows, uint64_t cols, float scale) {
  return guarded([&] {
    use(g);
    VKH_REQUIRE(kind == 1 || kind == 2, "loss_rows_grad: kind must be 1 (MSE) or 2 (Huber)");
    if (rows == 0) return;
    VKH_REQUIRE(x && y && L && dx, "loss_rows_grad: null pointer");
    unsigned grid = (unsigned)std::min<uint64_t>(ceil_div(rows, 256), (uint64_t)g->num_cus * 8);
    if (kind == 1)
      hipLaunchKernelGGL(k_loss_rows_grad<0>, dim3(grid), dim3(256), 0, g->stream, x, y, L, dx, rows, cols, scale);
    else
      hipLaunchKernelGGL(k_loss_rows_grad<1>, dim3(grid), dim3(256), 0, g->stream, x, y, L, dx, rows, cols, scale);
    check_launch("loss_rows_grad");
  });
}

int vkh_loss_grad(vkh_gpu* g, int kind, const float* x, const float* y, float* dx, uint64_t n, float scale) {
  return guarded([&] {
    use(g);
    if (n == 0) return;
    VKH_REQUIRE(x && y && dx, "loss_grad: null pointer");
    VKH_REQUIRE(kind >= 0 && kind <= 3, "loss_grad: unknown kind");
    hipLaunchKernelGGL(k_loss_grad, dim3(grid_for(g, n)), dim3(256), 0, g->stream, kind, x, y, dx, n, scale);
    check_launch("loss_grad");
  });
}

// 16-byte element loops when every pointer is 16-byte aligned
static uint64_t vec4_count(uint64_t n, std::initializer_list<const void*> ptrs) {
  for (const void* p : ptrs)
    if (p && ((uintptr_t)p & 15u)) return 0;
  return n / 4;
}

static RegTerms reg_terms(int kind, float l1, float s2) {
  VKH_REQUIRE(kind >= 0 && kind <= 3, "unknown regularizer kind");
  return RegTerms{kind, l1, s2};
}

int vkh_regularizer_grad(vkh_gpu* g, int kind, const float* w, float* grad, uint64_t n, float l1, float s2,
                         int accumulate) {
  return guarded([&] {
    use(g);
    VKH_REQUIRE(kind >= 1 && kind <= 3, "regularizer_grad: unknown regularizer kind");
    if (n == 0) return;
    VKH_REQUIRE(w && grad, "regularizer_grad: null pointer");
    const uint64_t n4 = vec4_count(n, {w, grad});
    hipLaunchKernelGGL(k_reg_grad, dim3(grid_for(g, n4 ? n4 : n)), dim3(256), 0, g->stream,
                       RegTerms{kind, l1, s2}, w, grad, n, n4, accumulate);
    check_launch("regularizer_grad");
  });
}

int vkh_adam(vkh_gpu* g, float* w, const float* gr, float* m, float* v, uint64_t n, float lr,
             double beta1, double beta2, float eps, double b1t, double b2t) {
  return guarded([&] {
    use(g);
    if (n == 0) return;
    VKH_REQUIRE(w && gr && m && v, "adam: null pointer");
    // (1 - beta) in FP64, then rounded: the Python float expression that the
    // reference passes to the FP32 scalar kernels
    AdamScalars s{(float)beta1, (float)(1.0 - beta1), (float)beta2, (float)(1.0 - beta2),
                  eps, -lr, (float)(1.0 - b1t), (float)(1.0 - b2t)};
    const uint64_t n4 = vec4_count(n, {w, gr, m, v});
    hipLaunchKernelGGL(k_adam, dim3(grid_for(g, n4 ? n4 : n)), dim3(256), 0, g->stream, w,
                       const_cast<float*>(gr), m, v, n, n4, s, (const double*)nullptr, RegTerms{0, 0.f, 0.f},
                       (uint16_t*)nullptr);
    check_launch("adam");
  });
}

int vkh_adam_pow_advance(vkh_gpu* g, double* bpow, double beta1, double beta2) {
  return guarded([&] {
    use(g);
    VKH_REQUIRE(bpow, "adam_pow_advance: null pointer");
    hipLaunchKernelGGL(k_adam_pow, dim3(1), dim3(64), 0, g->stream, bpow, beta1, beta2);
    check_launch("adam_pow_advance");
  });
}

int vkh_adam_device(vkh_gpu* g, float* w, float* gr, float* m, float* v, uint64_t n, float lr,
                    double beta1, double beta2, float eps, const double* bpow, uint16_t* w_bf16,
                    int reg_kind, float reg_l1, float reg_s2) {
  return guarded([&] {
    use(g);
    VKH_REQUIRE(w && gr && m && v && bpow, "adam_device: null pointer");
    const RegTerms r = reg_terms(reg_kind, reg_l1, reg_s2);
    if (n == 0) return;
    AdamScalars s{(float)beta1, (float)(1.0 - beta1), (float)beta2, (float)(1.0 - beta2),
                  eps, -lr, 1.0f, 1.0f};
    const uint64_t n4 = (w_bf16 && ((uintptr_t)w_bf16 & 7u)) ? 0 : vec4_count(n, {w, gr, m, v});
    hipLaunchKernelGGL(k_adam, dim3(grid_for(g, n4 ? n4 : n)), dim3(256), 0, g->stream, w, gr, m, v, n, n4,
                       s, bpow, r, w_bf16);
    check_launch("adam_device");
  });
}

int vkh_adam_device_multi(vkh_gpu* g, const vkh_adam_param* params, int count) {
  return guarded([&] {
    use(g);
    VKH_REQUIRE(count >= 0 && count <= ADAM_MT, "adam_device_multi: count out of range");
    VKH_REQUIRE(count == 0 || params, "adam_device_multi: null pointer");
    AdamMulti a{};
    uint64_t blocks = 0;
    int k = 0;
    for (int p = 0; p < count; ++p) {
      const vkh_adam_param& q = params[p];
      VKH_REQUIRE(q.w && q.g && q.m && q.v && q.bpow, "adam_device_multi: null pointer");
      const RegTerms r = reg_terms(q.reg_kind, q.reg_l1, q.reg_s2);
      if (q.n == 0) continue;
      const uint64_t n4 = vec4_count(q.n, {q.w, q.g, q.m, q.v});
      a.t[k] = AdamTensor{q.w, q.g, q.m, q.v, q.n, n4,
                          AdamScalars{(float)q.beta1, (float)(1.0 - q.beta1), (float)q.beta2,
                                      (float)(1.0 - q.beta2), q.eps, -q.lr, 1.0f, 1.0f},
                          q.bpow, r};
      a.first[k] = (uint32_t)blocks;
      blocks += grid_for(g, n4 ? n4 : q.n);
      ++k;
    }
    if (k == 0) return;
    VKH_REQUIRE(blocks < (uint64_t(1) << 31), "adam_device_multi: too many blocks");
    a.first[k] = (uint32_t)blocks;
    for (int p = k + 1; p <= ADAM_MT; ++p) a.first[p] = (uint32_t)blocks;
    a.count = k;
    hipLaunchKernelGGL(k_adam_multi, dim3((unsigned)blocks), dim3(256), 0, g->stream, a);
    check_launch("adam_device_multi");
  });
}

int vkh_sgd(vkh_gpu* g, float* w, float* gr, uint64_t n, float lr, int reg_kind, float reg_l1,
            float reg_s2) {
  return guarded([&] {
    use(g);
    const RegTerms r = reg_terms(reg_kind, reg_l1, reg_s2);
    if (n == 0) return;
    VKH_REQUIRE(w && gr, "sgd: null pointer");
    const uint64_t n4 = vec4_count(n, {w, gr});
    hipLaunchKernelGGL(k_sgd, dim3(grid_for(g, n4 ? n4 : n)), dim3(256), 0, g->stream, w, gr, n, n4, -lr, r);
    check_launch("sgd");
  });
}

int vkh_adagrad(vkh_gpu* g, float* w, float* gr, float* h, uint64_t n, float lr, float eps, int reg_kind,
                float reg_l1, float reg_s2) {
  return guarded([&] {
    use(g);
    const RegTerms r = reg_terms(reg_kind, reg_l1, reg_s2);
    if (n == 0) return;
    VKH_REQUIRE(w && gr && h, "adagrad: null pointer");
    const uint64_t n4 = vec4_count(n, {w, gr, h});
    hipLaunchKernelGGL(k_adagrad, dim3(grid_for(g, n4 ? n4 : n)), dim3(256), 0, g->stream, w, gr, h, n, n4,
                       -lr, eps, r);
    check_launch("adagrad");
  });
}

int vkh_bias_grad(vkh_gpu* g, const float* dy, float* db, uint64_t rows, uint64_t cols, float beta) {
  return guarded([&] {
    use(g);
    if (cols == 0) return;
    VKH_REQUIRE(dy && db, "bias_grad: null pointer");
    if (beta == 0.0f) {  // db = t exactly: the sums go straight to db
      int r = vkh_reduce_axis(g, VKH_RSUM, dy, db, 1, rows, cols, 0, 1.0f, 1.0f);
      if (r != VKH_OK) fail(r, vkh_last_error());
      return;
    }
    Workspace ws(g, sizeof(float) * cols);
    float* t = (float*)ws.get();
    int r = vkh_reduce_axis(g, VKH_RSUM, dy, t, 1, rows, cols, 0, 1.0f, 1.0f);
    if (r != VKH_OK) fail(r, vkh_last_error());
    hipLaunchKernelGGL(k_axpby, dim3(grid_for(g, cols)), dim3(256), 0, g->stream, db, (const float*)t, cols, beta);
    check_launch("bias_grad");
  });
}

}  // extern "C"
