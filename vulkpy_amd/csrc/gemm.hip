// GEMM on CDNA4 matrix cores (F15 matmul, F16 Dense affine).
//
// The reference computes every output with one thread looping over K from
// global memory (vulkpy/shader/matmul.comp:23-33, batch_affine.comp:25-39).
// Here a 256-thread block computes a BM x BN tile: operands are staged through
// LDS (K-contiguous rows, padded against bank conflicts), the next K slice is
// prefetched into registers while the current one is multiplied, and four
// waves (2 x 2) each own a (BM/2) x (BN/2) sub-tile built from 32 x 32 MFMA
// blocks:
//   FP32    v_mfma_f32_32x32x2_f32    (operands as stored)
//   BF16    v_mfma_f32_32x32x16_bf16  (operands rounded to nearest-even BF16)
//   BF16X3  the same instruction three times: hi*hi + hi*lo + lo*hi, where
//           x = hi + lo + O(2^-16 |x|) with hi = bf16(x), lo = bf16(x - hi)
// Accumulation is FP32 in every mode.  The epilogue applies alpha, beta*C,
// bias[n] and an activation, and can also emit a BF16 copy of the result.
// Blocks are remapped so that consecutive tiles land on the same XCD (each of
// the 8 XCDs has its own L2) and grouped along M for L2 reuse of B.
#include "common.h"

namespace vkh {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8_t;
typedef __attribute__((ext_vector_type(16))) float f32x16_t;

__device__ __forceinline__ uint16_t f2bf(float x) {
  // round to nearest even (NaN kept quiet)
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

constexpr int GEMM_THREADS = 256;

template <int MODE>
struct Cfg;
template <>
struct Cfg<VKH_GEMM_FP32> {
  static constexpr int BM = 128, BN = 128, BK = 16, PAD = 4;
  using T = float;  // LDS element
};
template <>
struct Cfg<VKH_GEMM_BF16> {
  static constexpr int BM = 128, BN = 128, BK = 32, PAD = 8;
  using T = uint16_t;
};
template <>
struct Cfg<VKH_GEMM_BF16X3> {
  static constexpr int BM = 128, BN = 128, BK = 32, PAD = 8;
  using T = uint16_t;
};

struct GemmArgs {
  const float* A;
  const float* B;
  float* C;
  const float* bias;
  uint16_t* Cbf;
  uint64_t M, N, K, lda, ldb, ldc;
  float alpha, beta;
  int act;
  const float* aux;  // activation output whose derivative multiplies the result
  uint64_t ldaux;
  int aux_act;
  uint32_t tiles_m, tiles_n;
};

__device__ __forceinline__ float act_apply(int act, float v) {
  if (act == VKH_ACT_RELU) return (v >= 0.0f || v != v) ? v : 0.0f;  // np.maximum(v, 0)
  if (act == VKH_ACT_SIGMOID) return 1.0f / (1.0f + (float)exp((double)(-v)));
  return v;
}

// dy * act'(y) given the activation's output y, in the op order of
// nn/layers.py:207-210 (ReLU) and 264-267 (Sigmoid)
__device__ __forceinline__ float act_deriv_mul(int act, float y, float dy) {
  if (act == VKH_ACT_RELU) {
    const float s = y > 0.f ? 1.f : (y < 0.f ? -1.f : (y == 0.f ? 0.f : y));
    const float d = (s >= 0.f || s != s) ? s : 0.f;
    return d * dy;
  }
  if (act == VKH_ACT_SIGMOID) {
    float d = 1.0f - y;
    d = d * y;
    return d * dy;
  }
  return dy;
}

// Reads 4 consecutive elements of a row (guarded at the row end).
__device__ __forceinline__ float4 load_row4(const float* row, uint64_t c, uint64_t limit, bool vec) {
  if (vec && c + 3 < limit) return *reinterpret_cast<const float4*>(row + c);
  float4 v;
  v.x = c < limit ? row[c] : 0.f;
  v.y = c + 1 < limit ? row[c + 1] : 0.f;
  v.z = c + 2 < limit ? row[c + 2] : 0.f;
  v.w = c + 3 < limit ? row[c + 3] : 0.f;
  return v;
}

// Staging of one operand tile (ROWS x BK) into LDS laid out [ROWS][BK+PAD]
// with K contiguous.  KMAJOR: the operand is stored with K contiguous
// (A row-major [M,K], or B stored [N,K]); otherwise rows are contiguous along
// the tile's row dimension (A stored [K,M], or B row-major [K,N]).
template <int MODE, int ROWS, bool KMAJOR>
struct Stager {
  using C = Cfg<MODE>;
  static constexpr int BK = C::BK;
  static constexpr int NV = ROWS * BK / 4 / GEMM_THREADS;  // float4 per thread
  float4 r[NV];

  __device__ __forceinline__ void load(const float* P, uint64_t ld, uint64_t row0, uint64_t nrows,
                                       uint64_t k0, uint64_t K, bool vec) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const int idx = threadIdx.x + i * GEMM_THREADS;
      if (KMAJOR) {
        const int row = idx / (BK / 4), c4 = idx % (BK / 4);
        const uint64_t gr = row0 + row;
        if (gr < nrows) r[i] = load_row4(P + gr * ld, k0 + c4 * 4, K, vec);
        else r[i] = make_float4(0.f, 0.f, 0.f, 0.f);
      } else {
        const int kr = idx / (ROWS / 4), c4 = idx % (ROWS / 4);
        const uint64_t gk = k0 + kr;
        if (gk < K) r[i] = load_row4(P + gk * ld, row0 + c4 * 4, nrows, vec);
        else r[i] = make_float4(0.f, 0.f, 0.f, 0.f);
      }
    }
  }

  // hi (and lo for split mode) tiles
  __device__ __forceinline__ void store(typename C::T* hi, typename C::T* lo) {
    constexpr int LD = BK + C::PAD;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const int idx = threadIdx.x + i * GEMM_THREADS;
      const float v[4] = {r[i].x, r[i].y, r[i].z, r[i].w};
      int base, step;
      if (KMAJOR) {
        const int row = idx / (BK / 4), c4 = idx % (BK / 4);
        base = row * LD + c4 * 4;
        step = 1;
      } else {
        const int kr = idx / (ROWS / 4), c4 = idx % (ROWS / 4);
        base = (c4 * 4) * LD + kr;
        step = LD;
      }
      if constexpr (MODE == VKH_GEMM_FP32) {
        if (KMAJOR) *reinterpret_cast<float4*>(&hi[base]) = r[i];
        else
#pragma unroll
          for (int q = 0; q < 4; ++q) hi[base + q * step] = v[q];
      } else {
        uint16_t h[4], l[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          h[q] = f2bf(v[q]);
          if constexpr (MODE == VKH_GEMM_BF16X3) l[q] = f2bf(v[q] - bf2f(h[q]));
        }
        if (KMAJOR) {
          uint2 hv = make_uint2((uint32_t)h[0] | ((uint32_t)h[1] << 16), (uint32_t)h[2] | ((uint32_t)h[3] << 16));
          *reinterpret_cast<uint2*>(&hi[base]) = hv;
          if constexpr (MODE == VKH_GEMM_BF16X3) {
            uint2 lv = make_uint2((uint32_t)l[0] | ((uint32_t)l[1] << 16), (uint32_t)l[2] | ((uint32_t)l[3] << 16));
            *reinterpret_cast<uint2*>(&lo[base]) = lv;
          }
        } else {
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            hi[base + q * step] = h[q];
            if constexpr (MODE == VKH_GEMM_BF16X3) lo[base + q * step] = l[q];
          }
        }
      }
    }
  }
};

// XCD-aware tile order: consecutive remapped ids share an XCD, then tiles are
// grouped by GROUP_M rows of tiles.
__device__ __forceinline__ void tile_coords(const GemmArgs& p, uint32_t& tm, uint32_t& tn) {
  const uint32_t nb = p.tiles_m * p.tiles_n;
  uint32_t b = blockIdx.x;
  if (nb % 8 == 0) b = (b % 8) * (nb / 8) + b / 8;
  constexpr uint32_t GROUP_M = 8;
  const uint32_t width = GROUP_M * p.tiles_n;
  const uint32_t group = b / width;
  const uint32_t first_m = group * GROUP_M;
  const uint32_t gsize = min(p.tiles_m - first_m, GROUP_M);
  tm = first_m + (b % width) % gsize;
  tn = (b % width) / gsize;
}

template <int MODE, bool TA, bool TB>
__global__ void __launch_bounds__(GEMM_THREADS, 2)
k_gemm(GemmArgs p, int vec_a, int vec_b) {
  using C = Cfg<MODE>;
  using T = typename C::T;
  constexpr int BM = C::BM, BN = C::BN, BK = C::BK, LD = BK + C::PAD;
  constexpr bool SPLIT = MODE == VKH_GEMM_BF16X3;
  __shared__ __attribute__((aligned(16))) T As[BM * LD];
  __shared__ __attribute__((aligned(16))) T Bs[BN * LD];
  __shared__ __attribute__((aligned(16))) T Asl[SPLIT ? BM * LD : 1];
  __shared__ __attribute__((aligned(16))) T Bsl[SPLIT ? BN * LD : 1];

  uint32_t tm, tn;
  tile_coords(p, tm, tn);
  const uint64_t m0 = (uint64_t)tm * BM, n0 = (uint64_t)tn * BN;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1;  // 2 x 2 waves, 64 x 64 each
  const int l32 = lane & 31, lhi = lane >> 5;

  Stager<MODE, BM, !TA> sa;
  Stager<MODE, BN, TB> sb;

  f32x16_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;

  const uint64_t K = p.K;
  sa.load(p.A, p.lda, m0, p.M, 0, K, vec_a);
  sb.load(p.B, p.ldb, n0, p.N, 0, K, vec_b);

  for (uint64_t k0 = 0; k0 < K; k0 += BK) {
    __syncthreads();  // previous slice fully consumed
    sa.store(As, Asl);
    sb.store(Bs, Bsl);
    __syncthreads();
    if (k0 + BK < K) {  // prefetch the next slice while the MFMAs run
      sa.load(p.A, p.lda, m0, p.M, k0 + BK, K, vec_a);
      sb.load(p.B, p.ldb, n0, p.N, k0 + BK, K, vec_b);
    }
    if constexpr (MODE == VKH_GEMM_FP32) {
#pragma unroll
      for (int ks = 0; ks < BK; ks += 2) {
        float a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) a[i] = As[(wm * 64 + i * 32 + l32) * LD + ks + lhi];
#pragma unroll
        for (int j = 0; j < 2; ++j) b[j] = Bs[(wn * 64 + j * 32 + l32) * LD + ks + lhi];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[i], b[j], acc[i][j], 0, 0, 0);
      }
    } else {
#pragma unroll
      for (int ks = 0; ks < BK; ks += 16) {
        bf16x8_t a[2], b[2], al[2], bl[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          const int off = (wm * 64 + i * 32 + l32) * LD + ks + lhi * 8;
          a[i] = *reinterpret_cast<const bf16x8_t*>(&As[off]);
          if constexpr (SPLIT) al[i] = *reinterpret_cast<const bf16x8_t*>(&Asl[off]);
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          const int off = (wn * 64 + j * 32 + l32) * LD + ks + lhi * 8;
          b[j] = *reinterpret_cast<const bf16x8_t*>(&Bs[off]);
          if constexpr (SPLIT) bl[j] = *reinterpret_cast<const bf16x8_t*>(&Bsl[off]);
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            if constexpr (SPLIT) {
              acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], b[j], acc[i][j], 0, 0, 0);
              acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], bl[j], acc[i][j], 0, 0, 0);
            }
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
          }
      }
    }
  }

  // epilogue: acc[i][j][r] is C[m, n] with
  //   m = m0 + wm*64 + i*32 + (r/4)*8 + lhi*4 + r%4,  n = n0 + wn*64 + j*32 + l32
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const uint64_t n = n0 + wn * 64 + j * 32 + l32;
      if (n >= p.N) continue;
      const float bn = p.bias ? p.bias[n] : 0.f;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const uint64_t m = m0 + wm * 64 + i * 32 + (r / 4) * 8 + lhi * 4 + (r % 4);
        if (m >= p.M) continue;
        float v = p.alpha * acc[i][j][r];
        float* cp = p.C + m * p.ldc + n;
        if (p.beta != 0.f) v += p.beta * *cp;
        if (p.bias) v += bn;
        v = act_apply(p.act, v);
        if (p.aux) v = act_deriv_mul(p.aux_act, p.aux[m * p.ldaux + n], v);
        *cp = v;
        if (p.Cbf) p.Cbf[m * p.ldc + n] = f2bf(v);
      }
    }
}

template <int MODE>
static void launch_mode(vkh_gpu* g, bool ta, bool tb, GemmArgs& a, int va, int vb) {
  using C = Cfg<MODE>;
  a.tiles_m = (uint32_t)ceil_div(a.M, C::BM);
  a.tiles_n = (uint32_t)ceil_div(a.N, C::BN);
  const uint64_t nb = (uint64_t)a.tiles_m * a.tiles_n;
  VKH_REQUIRE(nb < (uint64_t(1) << 31), "GEMM too large");
  dim3 grid((unsigned)nb), block(GEMM_THREADS);
  if (!ta && !tb) hipLaunchKernelGGL((k_gemm<MODE, false, false>), grid, block, 0, g->stream, a, va, vb);
  else if (!ta && tb) hipLaunchKernelGGL((k_gemm<MODE, false, true>), grid, block, 0, g->stream, a, va, vb);
  else if (ta && !tb) hipLaunchKernelGGL((k_gemm<MODE, true, false>), grid, block, 0, g->stream, a, va, vb);
  else hipLaunchKernelGGL((k_gemm<MODE, true, true>), grid, block, 0, g->stream, a, va, vb);
  check_launch("gemm");
}

// ---- helpers for the nn path -------------------------------------------------------------
__global__ void __launch_bounds__(256) k_to_bf16(const float* src, uint16_t* dst, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) dst[i] = f2bf(src[i]);
}

}  // namespace vkh

using namespace vkh;

extern "C" {

static void gemm_impl(vkh_gpu* g, int mode, int trans_a, int trans_b, uint64_t M, uint64_t N,
                      uint64_t K, float alpha, const float* A, uint64_t lda, const float* B,
                      uint64_t ldb, float beta, float* Cp, uint64_t ldc, const float* bias, int act,
                      uint16_t* c_bf16, const float* aux, uint64_t ldaux, int aux_act) {
  use(g);
  if (M == 0 || N == 0) return;
  VKH_REQUIRE(Cp, "gemm: null C");
  VKH_REQUIRE(K == 0 || (A && B), "gemm: null operand");
  VKH_REQUIRE(ldc >= N, "gemm: ldc < N");
  VKH_REQUIRE(trans_a ? lda >= M : lda >= K, "gemm: lda too small");
  VKH_REQUIRE(trans_b ? ldb >= K : ldb >= N, "gemm: ldb too small");
  VKH_REQUIRE(act >= VKH_ACT_NONE && act <= VKH_ACT_SIGMOID, "gemm: unknown activation");
  VKH_REQUIRE(aux_act >= VKH_ACT_NONE && aux_act <= VKH_ACT_SIGMOID, "gemm: unknown aux activation");
  VKH_REQUIRE(!aux || ldaux >= N, "gemm: ldaux < N");
  GemmArgs a{};
  a.A = A; a.B = B; a.C = Cp; a.bias = bias; a.Cbf = c_bf16;
  a.M = M; a.N = N; a.K = K; a.lda = lda; a.ldb = ldb; a.ldc = ldc;
  a.alpha = alpha; a.beta = beta; a.act = act;
  a.aux = aux; a.ldaux = ldaux; a.aux_act = aux_act;
  const int va = ((uintptr_t)A & 15u) == 0 && lda % 4 == 0;
  const int vb = ((uintptr_t)B & 15u) == 0 && ldb % 4 == 0;
  switch (mode) {
    case VKH_GEMM_FP32: launch_mode<VKH_GEMM_FP32>(g, trans_a, trans_b, a, va, vb); break;
    case VKH_GEMM_BF16: launch_mode<VKH_GEMM_BF16>(g, trans_a, trans_b, a, va, vb); break;
    case VKH_GEMM_BF16X3: launch_mode<VKH_GEMM_BF16X3>(g, trans_a, trans_b, a, va, vb); break;
    default: fail(VKH_INVALID, "gemm: unknown mode");
  }
}

int vkh_gemm(vkh_gpu* g, int mode, int trans_a, int trans_b, uint64_t M, uint64_t N, uint64_t K,
             float alpha, const float* A, uint64_t lda, const float* B, uint64_t ldb, float beta,
             float* Cp, uint64_t ldc, const float* bias, int act, uint16_t* c_bf16) {
  return guarded([&] {
    gemm_impl(g, mode, trans_a, trans_b, M, N, K, alpha, A, lda, B, ldb, beta, Cp, ldc, bias, act,
              c_bf16, nullptr, 0, VKH_ACT_NONE);
  });
}

int vkh_gemm_ex(vkh_gpu* g, int mode, int trans_a, int trans_b, uint64_t M, uint64_t N, uint64_t K,
                float alpha, const float* A, uint64_t lda, const float* B, uint64_t ldb, float beta,
                float* Cp, uint64_t ldc, const float* bias, int act, uint16_t* c_bf16,
                const float* aux, uint64_t ldaux, int aux_act) {
  return guarded([&] {
    gemm_impl(g, mode, trans_a, trans_b, M, N, K, alpha, A, lda, B, ldb, beta, Cp, ldc, bias, act,
              c_bf16, aux, ldaux, aux_act);
  });
}

int vkh_to_bf16(vkh_gpu* g, const float* src, uint16_t* dst, uint64_t n) {
  return guarded([&] {
    use(g);
    if (n == 0) return;
    VKH_REQUIRE(src && dst, "to_bf16: null pointer");
    unsigned grid = stream_grid(g, n, 256, 8);
    hipLaunchKernelGGL(k_to_bf16, dim3(grid), dim3(256), 0, g->stream, src, dst, n);
    check_launch("to_bf16");
  });
}

}  // extern "C"
